// The reference program's relax protocol on top of the C ABI:
// parse_batch_file (large_strain_relax.cu:297-383), get_file_lists
// (file_lists.cpp:80-137), do_relaxation_run (:792-1010) and
// relaxation_batch (:1016-1044), with the same output files and formats.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <memory>
#include <filesystem>
#include <fstream>
#include <regex>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/nrelax.h"
#include "internal.hpp"
#include "protocol.hpp"

namespace nr {

namespace {

constexpr double kFltTol = 1e-10;

// read_line :263-283: skip blank and '#' lines.
bool read_line(std::ifstream &in, std::string &line)
{
    if (in.eof()) return false;
    while (!in.eof()) {
        std::getline(in, line);
        if (line.empty() || line[0] == '#') continue;
        break;
    }
    return !(line.empty() || line[0] == '#');
}

std::vector<std::string> split_ws(const std::string &s)
{
    std::vector<std::string> out;
    std::string cur;
    for (char c : s) {
        if (c == ' ' || c == '\t') {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(c);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

}  // namespace

bool valid_range(double mn, double mx, double inc)
{
    if (!std::isfinite(mn) || !std::isfinite(mx) || !std::isfinite(inc)) return false;
    if (mn < 0 || mx < 0 || inc < 0 || mx < mn) return false;
    // a ramp of more increments than this is a mistake, not a protocol
    return !(inc > 0) || (mx - mn) / inc <= kMaxRampSteps;
}

int parse_batch_file(const std::string &path, BatchSpec &b, std::string &err)
{
    std::ifstream in(path);
    std::string line;
    b = BatchSpec();
    if (!in.is_open()) {
        err = "The file " + path + " could not be read.";
        return NR_E_OPEN;
    }
    if (!read_line(in, line)) { err = "Not enough information was provided in the batch file."; return NR_E_BATCH; }
    std::vector<std::string> tok = split_ws(line);
    if (tok.empty()) { err = "missing base name"; return NR_E_BATCH; }
    b.base = tok[0];
    if (tok.size() > 1) b.tag = tok[1];

    if (!read_line(in, line)) { err = "Not enough information was provided in the batch file."; return NR_E_BATCH; }
    double klo = 0, khi = 0, seed = 0;
    int nr = std::sscanf(line.c_str(), "%lf %lf %lf %lf %lf", &b.ks, &b.kb, &klo, &khi, &seed);
    if (nr < 2 || !(b.ks >= 0) || !(b.kb >= 0) || !std::isfinite(b.ks) || !std::isfinite(b.kb)) { err = "Error: invalid format for mechanical attributes."; return NR_E_BATCH; }
    if (nr == 5 && klo > 0 && khi >= klo) {
        b.kappa_lo = klo;
        b.kappa_hi = khi;
        b.kappa_seed = (uint64_t)seed;
    }

    if (!read_line(in, line)) { err = "Not enough information was provided in the batch file."; return NR_E_BATCH; }
    int cnt = 0;
    for (const std::string &t : split_ws(line)) {
        double x;
        if (cnt < 5 && std::sscanf(t.c_str(), "%lf", &x) == 1) {
            if (!std::isfinite(x)) { err = "FIRE parameters must be finite."; return NR_E_BATCH; }
            b.fp[cnt++] = x;
        }
    }
    if (cnt < 5) { err = "Too few FIRE parameters were supplied."; return NR_E_BATCH; }

    if (!read_line(in, line)) { err = "Not enough information was provided in the batch file."; return NR_E_BATCH; }
    char w1[64] = {0}, w2[64] = {0};
    nr = std::sscanf(line.c_str(), "%d %lf %lf %d %63s %63s", &b.nmin, &b.mass, &b.fcut, &b.nmax, w1, w2);
    if (nr < 4 || !std::isfinite(b.mass) || !std::isfinite(b.fcut)) { err = "One or more of nmin, mass, cutoff force and step limit were invalid."; return NR_E_BATCH; }
    if (!std::strcmp(w1, "max") || !std::strcmp(w2, "max")) b.criterion = NR_CRIT_MAX;
    if (!std::strcmp(w1, "cg") || !std::strcmp(w2, "cg")) b.method = NR_METHOD_CG;
    if (!std::strcmp(w1, "cgpr") || !std::strcmp(w2, "cgpr")) b.method = NR_METHOD_CG_PR;

    double mn, mx, inc;
    if (!read_line(in, line)) { err = "Not enough information was provided in the batch file."; return NR_E_BATCH; }
    nr = std::sscanf(line.c_str(), "%lf %lf %lf", &mn, &mx, &inc);
    if (nr < 3 || !valid_range(mn, mx, inc)) { err = "Error: invalid specification of compression."; return NR_E_BATCH; }
    b.cmin = mn; b.cmax = mx; b.cinc = inc;

    if (!read_line(in, line)) { err = "Not enough information was provided in the batch file."; return NR_E_BATCH; }
    nr = std::sscanf(line.c_str(), "%lf %lf %lf", &mn, &mx, &inc);
    if (nr < 3 || !valid_range(mn, mx, inc)) { err = "Error: invalid specification of shear."; return NR_E_BATCH; }
    b.smin = mn; b.smax = mx; b.sinc = inc;
    return NR_OK;
}

// get_file_lists, file_lists.cpp:80-137: files of `dir` matching
// ^(<base>.*)\.<ext>; outputs <dir><m>_<tag>_log.txt, _edata.txt and the
// displacement base <dir><m>_<tag>.  The input is split into '/' tokens
// (empty tokens dropped, :93-103): the last token is the base, and the others
// rebuild the directory (a leading '/' kept) only when there are at least two
// tokens; otherwise the directory is "./" ("/x" searches ./ for x*, "dir/"
// searches ./ for dir*).  The four names of one network are kept together
// (the reference sorts the four lists independently, which can pair one
// network's input with another's outputs: net1 and net10).
int file_lists(const std::string &input, const std::string &tag, std::vector<JobFiles> &jobs,
               std::string &err, const std::string &ext)
{
    jobs.clear();
    if (input.empty()) { err = "empty base name"; return NR_E_BATCH; }
    std::vector<std::string> toks;
    std::string cur;
    for (char c : input) {
        if (c == '/') {
            if (!cur.empty()) toks.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(c);
        }
    }
    if (!cur.empty()) toks.push_back(cur);
    if (toks.empty()) { err = "empty base name"; return NR_E_BATCH; }
    std::string base = toks.back(), dir = "./";
    if (toks.size() > 1) {
        std::ostringstream oss;
        if (input[0] == '/') oss << "/";
        for (size_t i = 0; i + 1 < toks.size(); i++) oss << toks[i] << "/";
        dir = oss.str();
    }
    if (base == "*") base = "";
    std::regex expr;
    try {
        expr = std::regex("^(" + base + ".*)\\." + ext);
    } catch (const std::regex_error &) {
        err = "base name is not a valid pattern: " + base;
        return NR_E_BATCH;
    }
    std::error_code ec;
    std::filesystem::path p(dir);
    if (!std::filesystem::exists(p, ec)) return NR_OK;
    std::vector<std::string> names;
    std::filesystem::directory_iterator it(p, ec), end;
    for (; !ec && it != end; it.increment(ec)) {
        std::error_code ed;
        if (it->is_directory(ed)) continue;
        names.push_back(it->path().filename().string());
    }
    if (ec) {
        err = "directory " + dir + " could not be listed: " + ec.message();
        return NR_E_OPEN;
    }
    std::sort(names.begin(), names.end());
    for (const std::string &name : names) {
        std::smatch m;
        bool hit = false;
        try {
            hit = std::regex_search(name, m, expr);
        } catch (const std::regex_error &) {
            err = "base name pattern too complex for " + name;
            return NR_E_BATCH;
        }
        if (!hit) continue;
        JobFiles j;
        std::string stem = m[1].str();
        std::string t = tag.empty() ? "" : "_" + tag;
        j.dat = dir + m[0].str();
        j.log = dir + stem + t + "_log.txt";
        j.edata = dir + stem + t + "_edata.txt";
        j.disp = dir + stem + t;
        jobs.push_back(j);
    }
    std::sort(jobs.begin(), jobs.end(), [](const JobFiles &a, const JobFiles &b) { return a.dat < b.dat; });
    return NR_OK;
}

int dec_digits(int value)
{
    int d = 0;
    do {
        d++;
        value /= 10;
    } while (value > 0);
    return d;
}

// report_name :203-213
std::string report_name(const std::string &base, int digits, int count, const std::string &ext)
{
    std::ostringstream oss;
    oss << base << "_";
    for (int k = 0; k < digits - dec_digits(count); k++) oss << 0;
    oss << count << "." << ext;
    return oss.str();
}

namespace {

void log_result(FILE *lf, const nr_report &r, double fcut, int criterion)
{
    const double measure = criterion == NR_CRIT_MAX ? r.max_force : r.rms_force;
    if (lf) {
        std::fprintf(lf, measure <= fcut ? "Convergence reached: " : "Convergence not reached: ");
        std::fprintf(lf, "Steps: %d, RMS Force: %12.8le\n", r.steps, r.rms_force);
    }
    std::printf("Steps: %d, RMS Force: %12.8le\n", r.steps, r.rms_force);
}

}  // namespace

// The set-up of do_relaxation_run (:816-902): read, triples, device arrays.
int proto_open(const JobFiles &jf, const BatchSpec &b, const nr_batch_opts *opts, ProtoNet &pn,
               std::string &err)
{
    int rc = nr_network_read(jf.dat.c_str(), &pn.n);
    if (rc) {
        err = nr_last_error();
        return rc;
    }
    if (b.kb > 0) {
        nr_network_build_triples(pn.n);
        if (b.kappa_hi > 0) nr_network_set_fiber_kappa(pn.n, b.kappa_lo, b.kappa_hi, b.kappa_seed);
    }
    nr_network_get_info(pn.n, &pn.info);
    rc = nr_batch_create(1, &pn.n, &b.ks, &b.kb, opts, &pn.b);
    if (rc) err = nr_last_error();
    return rc;
}

// The compression then shear ramps of do_relaxation_run (:917-1000) from the
// network's resident state, each increment an affine step of `inc` (the first
// one included) and a relax; the log, edata and .pos outputs, each when its
// flag is set.
int proto_ramps(ProtoNet &pn, const JobFiles &jf, const BatchSpec &b, bool strain_mode,
                unsigned report, std::string &err)
{
    std::clock_t ct = std::clock();
    const nr_network_info &info = pn.info;
    const double range = info.height > 0 ? info.height : info.maxy - info.miny;
    const double smult = strain_mode ? range : 1.0;
    if ((b.cmax > 0 && !(b.cinc > 0)) || (b.smax > 0 && !(b.sinc > 0))) {
        err = "a strain ramp with a maximum above 0 needs a positive increment";
        return NR_E_BATCH;
    }
    if (info.height > 0 && b.cmax > 0) {
        err = "compression is not defined in the Lees-Edwards box";
        return NR_E_BATCH;
    }
    const int num_digits = dec_digits(1 + (b.sinc > 0 ? (int)((b.smax - b.smin) / b.sinc) : 0));
    nr_relax_params rp;
    rp.alpha0 = b.fp[0];
    rp.f_alpha = b.fp[1];
    rp.dt_max = b.fp[2];
    rp.f_inc = b.fp[3];
    rp.f_dec = b.fp[4];
    rp.nmin = b.nmin;
    rp.mass = b.mass;
    rp.fcut = b.fcut;
    rp.max_steps = b.nmax;
    rp.criterion = b.criterion;
    rp.method = b.method;

    FILE *lf = (report & kReportLog) ? std::fopen(jf.log.c_str(), "w") : nullptr;
    FILE *ef = (report & kReportEdata) ? std::fopen(jf.edata.c_str(), "w") : nullptr;
    std::vector<double> pos(2 * (size_t)info.n_nodes);
    double mn = b.cmin * smult, mx = b.cmax * smult, inc = b.cinc * smult, cur = 0.0;
    nr_report rep;
    int rc = 0;
    if (mx > 0) {
        for (cur = mn; cur <= mx + kFltTol; cur += inc) {
            if ((rc = nr_batch_strain(pn.b, 0, 0.0, -inc)) != 0 ||
                (rc = nr_batch_relax(pn.b, &rp, 1, &rep)) != 0) {
                err = nr_last_error();
                break;
            }
            log_result(lf, rep, b.fcut, b.criterion);
            if (ef)
                std::fprintf(ef, "%12.0lf\t%12.10lf\t%12.8le\t%12.8le\t%12.8le\n", 0., cur,
                             rep.e_stretch, rep.e_bend, rep.e_stretch + rep.e_bend);
        }
    }
    const double max_comp = mx > mn ? cur - inc : mx;
    mn = b.smin * smult;
    mx = b.smax * smult;
    inc = b.sinc * smult;
    int disp_count = 1;
    if (rc == 0 && mx > 0) {
        for (cur = mn; cur <= mx + kFltTol; cur += inc) {
            if ((rc = nr_batch_strain(pn.b, 0, inc, 0.0)) != 0 ||
                (rc = nr_batch_relax(pn.b, &rp, 1, &rep)) != 0 ||
                ((report & kReportDisp) &&
                 (rc = nr_batch_get_positions(pn.b, 0, pos.data())) != 0)) {
                err = nr_last_error();
                break;
            }
            log_result(lf, rep, b.fcut, b.criterion);
            if (ef)
                std::fprintf(ef, "%12.10lf\t%12.10lf\t%12.10le\t%12.10le\t%12.10le\n", cur, max_comp,
                             rep.e_stretch, rep.e_bend, rep.e_stretch + rep.e_bend);
            if (!(report & kReportDisp)) continue;
            std::string name = report_name(jf.disp, num_digits, disp_count++, "pos");
            FILE *pf = std::fopen(name.c_str(), "w");
            if (pf) {
                for (int k = 0; k < info.n_nodes; k++)
                    std::fprintf(pf, "%2.10le\t%2.10le\n", pos[2 * k], pos[2 * k + 1]);
                std::fclose(pf);
            }
        }
    }
    ct = std::clock() - ct;
    if (lf) {
        std::fprintf(lf, "Elapsed time: %f seconds\n", ((float)ct) / CLOCKS_PER_SEC);
        std::fclose(lf);
    }
    if (ef) std::fclose(ef);
    return rc;
}

// do_relaxation_run :792-1010 for one network.
int run_protocol(const JobFiles &jf, const BatchSpec &b, bool strain_mode,
                 const nr_batch_opts *opts, std::string &err)
{
    ProtoNet pn;
    int rc = proto_open(jf, b, opts, pn, err);
    if (rc) return rc;
    return proto_ramps(pn, jf, b, strain_mode, kReportAll, err);
}

namespace {

// Nodes of the networks one device batch of relaxation_batch holds at once
// (about 10 GB of device state at 64 M nodes).
constexpr int64_t kProtocolBatchNodes = int64_t(1) << 26;

// One network of a batched protocol run: its files, its ramp in its own
// displacement units, and where its loops stand.
struct BatchedNet {
    const JobFiles *jf = nullptr;
    nr_network *n = nullptr;
    nr_network_info info{};
    double range = 0, cmn = 0, cmx = 0, cinc = 0, smn = 0, smx = 0, sinc = 0, cur = 0;
    double max_comp = 0;
    int disp_count = 1;
    FILE *lf = nullptr, *ef = nullptr;
    bool active = false;
};

void close_net(BatchedNet &z, std::clock_t ct)
{
    if (z.lf) {
        std::fprintf(z.lf, "Elapsed time: %f seconds\n", ((float)ct) / CLOCKS_PER_SEC);
        std::fclose(z.lf);
    }
    if (z.ef) std::fclose(z.ef);
    z.lf = z.ef = nullptr;
}

// The compression then shear ramps of do_relaxation_run (:917-1000) for
// every network of `nets` at once: each round applies every active
// network's increment (scaled by its own y range in strain mode) and relaxes
// them all in one persistent launch; a network whose loop has ended sits out
// (step limit 0: its state is left as it is).  Outputs are each network's
// own, in the reference's formats.
int ramps_batched(std::vector<BatchedNet> &nets, nr_batch *bt, const BatchSpec &b,
                  std::string &err)
{
    const int count = (int)nets.size();
    nr_relax_params rp;
    rp.alpha0 = b.fp[0];
    rp.f_alpha = b.fp[1];
    rp.dt_max = b.fp[2];
    rp.f_inc = b.fp[3];
    rp.f_dec = b.fp[4];
    rp.nmin = b.nmin;
    rp.mass = b.mass;
    rp.fcut = b.fcut;
    rp.max_steps = b.nmax;
    rp.criterion = b.criterion;
    rp.method = b.method;
    std::vector<nr_relax_params> prm(count, rp);
    std::vector<nr_report> rep(count);
    std::vector<double> pos;
    for (int pass = 0; pass < 2; pass++) {  // compression, then shear
        const bool shear = pass == 1;
        for (BatchedNet &z : nets) {
            const double mn = shear ? z.smn : z.cmn, mx = shear ? z.smx : z.cmx;
            z.cur = mn;
            z.active = mx > 0 && z.cur <= mx + kFltTol;
        }
        for (;;) {
            bool any = false;
            for (int k = 0; k < count; k++) {
                BatchedNet &z = nets[k];
                prm[k].max_steps = z.active ? b.nmax : 0;
                if (!z.active) continue;
                any = true;
                const int rc = shear ? nr_batch_strain(bt, k, z.sinc, 0.0)
                                     : nr_batch_strain(bt, k, 0.0, -z.cinc);
                if (rc) {
                    err = nr_last_error();
                    return rc;
                }
            }
            if (!any) break;
            int rc = nr_batch_relax(bt, prm.data(), 0, rep.data());
            if (rc) {
                err = nr_last_error();
                return rc;
            }
            for (int k = 0; k < count; k++) {
                BatchedNet &z = nets[k];
                if (!z.active) continue;
                log_result(z.lf, rep[k], b.fcut, b.criterion);
                if (!shear) {
                    if (z.ef)
                        std::fprintf(z.ef, "%12.0lf\t%12.10lf\t%12.8le\t%12.8le\t%12.8le\n", 0.,
                                     z.cur, rep[k].e_stretch, rep[k].e_bend,
                                     rep[k].e_stretch + rep[k].e_bend);
                } else {
                    if (z.ef)
                        std::fprintf(z.ef, "%12.10lf\t%12.10lf\t%12.10le\t%12.10le\t%12.10le\n",
                                     z.cur, z.max_comp, rep[k].e_stretch, rep[k].e_bend,
                                     rep[k].e_stretch + rep[k].e_bend);
                    pos.resize(2 * (size_t)z.info.n_nodes);
                    if ((rc = nr_batch_get_positions(bt, k, pos.data())) != 0) {
                        err = nr_last_error();
                        return rc;
                    }
                    const int digits =
                        dec_digits(1 + (b.sinc > 0 ? (int)((b.smax - b.smin) / b.sinc) : 0));
                    const std::string name = report_name(z.jf->disp, digits, z.disp_count++, "pos");
                    if (FILE *pf = std::fopen(name.c_str(), "w")) {
                        for (int i = 0; i < z.info.n_nodes; i++)
                            std::fprintf(pf, "%2.10le\t%2.10le\n", pos[2 * i], pos[2 * i + 1]);
                        std::fclose(pf);
                    }
                }
                z.cur += shear ? z.sinc : z.cinc;
                z.active = z.cur <= (shear ? z.smx : z.cmx) + kFltTol;
            }
        }
        if (!shear)  // :957 max_comp from the loop's exit value
            for (BatchedNet &z : nets) z.max_comp = z.cmx > z.cmn ? z.cur - z.cinc : z.cmx;
    }
    return NR_OK;
}

}  // namespace

// relaxation_batch (:1016-1044) with every network of the directory in one
// resident device batch (in groups of at most kProtocolBatchNodes nodes):
// per network the files and numbers of do_relaxation_run, the increments of
// all networks relaxed together.  A network that cannot be read or set up is
// reported and skipped, as the reference moves on to the next file.  The
// elapsed time in a network's log is that of its group.  With a fixed
// tiles_per_member in opts each network's arithmetic, and so its files, do
// not depend on the other networks.
int run_protocol_batched(const std::vector<JobFiles> &jobs, const BatchSpec &b, bool strain_mode,
                         const nr_batch_opts *opts, int &first_err)
{
    first_err = NR_OK;
    auto note = [&](const JobFiles &j, int rc, const std::string &msg) {
        std::fprintf(stderr, "%s: %s\n", j.dat.c_str(), msg.c_str());
        if (!first_err) first_err = rc;
    };
    if ((b.cmax > 0 && !(b.cinc > 0)) || (b.smax > 0 && !(b.sinc > 0))) {
        for (const JobFiles &j : jobs)
            note(j, NR_E_BATCH, "a strain ramp with a maximum above 0 needs a positive increment");
        return first_err;
    }
    size_t next = 0;
    while (next < jobs.size()) {
        std::clock_t ct = std::clock();
        std::vector<BatchedNet> nets;
        std::vector<nr_network *> handles;
        int64_t nodes = 0;
        for (; next < jobs.size() && (nets.empty() || nodes < kProtocolBatchNodes); next++) {
            const JobFiles &j = jobs[next];
            BatchedNet z;
            z.jf = &j;
            int rc = nr_network_read(j.dat.c_str(), &z.n);
            if (rc) {
                note(j, rc, nr_last_error());
                continue;
            }
            if (b.kb > 0) {
                nr_network_build_triples(z.n);
                if (b.kappa_hi > 0)
                    nr_network_set_fiber_kappa(z.n, b.kappa_lo, b.kappa_hi, b.kappa_seed);
            }
            nr_network_get_info(z.n, &z.info);
            if (z.info.height > 0 && b.cmax > 0) {
                note(j, NR_E_BATCH, "compression is not defined in the Lees-Edwards box");
                nr_network_free(z.n);
                continue;
            }
            z.range = z.info.height > 0 ? z.info.height : z.info.maxy - z.info.miny;
            const double smult = strain_mode ? z.range : 1.0;
            z.cmn = b.cmin * smult;
            z.cmx = b.cmax * smult;
            z.cinc = b.cinc * smult;
            z.smn = b.smin * smult;
            z.smx = b.smax * smult;
            z.sinc = b.sinc * smult;
            nodes += z.info.n_nodes;
            handles.push_back(z.n);
            nets.push_back(z);
        }
        if (nets.empty()) continue;
        std::vector<double> ks(nets.size(), b.ks), kb(nets.size(), b.kb);
        nr_batch *bt = nullptr;
        int rc = nr_batch_create((int)nets.size(), handles.data(), ks.data(), kb.data(), opts, &bt);
        std::string err = rc ? nr_last_error() : "";
        if (!rc) {
            for (BatchedNet &z : nets) {
                z.lf = std::fopen(z.jf->log.c_str(), "w");
                z.ef = std::fopen(z.jf->edata.c_str(), "w");
            }
            rc = ramps_batched(nets, bt, b, err);
            nr_batch_destroy(bt);
        }
        ct = std::clock() - ct;
        for (BatchedNet &z : nets) {
            close_net(z, ct);
            if (rc) note(*z.jf, rc, err);
            nr_network_free(z.n);
        }
    }
    return first_err;
}

// parse_batch_file of sparse_relax.cpp (:920-1035).
int parse_linear_batch(const std::string &path, LinearBatchSpec &b, std::string &err)
{
    std::ifstream in(path);
    std::string line;
    b = LinearBatchSpec();
    if (!in.is_open()) {
        err = "The file " + path + " could not be read.";
        return NR_E_OPEN;
    }
    if (!read_line(in, line)) { err = "Not enough information was found."; return NR_E_BATCH; }
    std::vector<std::string> tok = split_ws(line);
    if (tok.empty()) { err = "Not enough information was found."; return NR_E_BATCH; }
    b.base = tok[0];
    if (tok.size() > 1) b.tag = tok[1];
    if (!read_line(in, line)) { err = "Not enough information was provided in the batch file."; return NR_E_BATCH; }
    int nr = std::sscanf(line.c_str(), "%lf %lf %lf", &b.ks, &b.kb, &b.mu);
    if (nr < 3 || !(b.ks >= 0) || !(b.kb >= 0) || !(b.mu >= 0) || !std::isfinite(b.ks) ||
        !std::isfinite(b.kb) || !std::isfinite(b.mu)) { err = "Error: invalid format for mechanical attributes."; return NR_E_BATCH; }
    double mn, mx, inc;
    if (!read_line(in, line)) { err = "Not enough information was provided in the batch file."; return NR_E_BATCH; }
    nr = std::sscanf(line.c_str(), "%lf %lf %lf", &mn, &mx, &inc);
    if (nr < 3 || !valid_range(mn, mx, inc)) { err = "Error: invalid specification of compression."; return NR_E_BATCH; }
    b.cmin = mn; b.cmax = mx; b.cinc = inc;
    if (!read_line(in, line)) { err = "Not enough information was provided in the batch file."; return NR_E_BATCH; }
    nr = std::sscanf(line.c_str(), "%lf %lf %lf", &mn, &mx, &inc);
    if (nr < 3 || !valid_range(mn, mx, inc)) { err = "Error: invalid specification of shear."; return NR_E_BATCH; }
    b.smin = mn; b.smax = mx; b.sinc = inc;
    return NR_OK;
}

// do_relaxation_run of sparse_relax.cpp (:676-881) for one network: for
// every compression, then shear, of the load grid, one small-strain solve
// from the input geometry; per load step a residual line in the log, an
// energy line and a displacement file.  In strain mode the compression is
// scaled by the y extent and the shear by the x extent (:731-739), then both
// are divided by the y extent to form the affine field (:785), as the
// reference does.  plan() reads the network and lists its solves (jobs), the
// caller solves them (with other networks' in one device batch), write()
// writes the outputs.
struct LinearRun {
    const JobFiles *jf = nullptr;
    std::time_t start = 0;
    nr_network *net = nullptr;
    int n_nodes = 0;
    bool itemize = false;
    std::vector<double> comps, shears;      // load steps, in the reference's order
    std::vector<std::vector<double>> u;     // per step: bc + w
    std::vector<nr_linear_job> jobs;
    std::vector<nr_linear_report> reps;
    int digits = 1;
    ~LinearRun() { nr_network_free(net); }

    int plan(const JobFiles &f, const LinearBatchSpec &spec, bool strain_mode, bool affine,
             bool item, std::string &err)
    {
        jf = &f;
        itemize = item;
        start = std::time(nullptr);
        int rc = nr_network_read(f.dat.c_str(), &net);
        if (rc) {
            err = nr_last_error();
            return rc;
        }
        if (spec.kb > 0) nr_network_build_triples(net);
        nr_network_info info;
        nr_network_get_info(net, &info);
        n_nodes = info.n_nodes;
        std::vector<double> pos(2 * (size_t)info.n_nodes);
        nr_network_get_arrays(net, pos.data(), nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                              nullptr);
        double minx = DBL_MAX, maxx = -DBL_MAX;
        for (int i = 0; i < info.n_nodes; i++) {
            minx = std::min(minx, pos[2 * i]);
            maxx = std::max(maxx, pos[2 * i]);
        }
        const double yext = info.maxy - info.miny, xext = maxx - minx;
        // equal ends: one step (run_relaxation_batch :1081-1082), before scaling
        double cmin = spec.cmin, cmax = spec.cmax, cinc = spec.cinc;
        double smin = spec.smin, smax = spec.smax, sinc = spec.sinc;
        if (smax - smin < kFltTol) sinc = 1e12;
        if (cmax - cmin < kFltTol) cinc = 1e12;
        if (strain_mode) {
            cmin *= yext;
            cmax *= yext;
            cinc *= yext;
            smin *= xext;
            smax *= xext;
            sinc *= xext;
        }
        if (!(cinc > 0) || !(sinc > 0)) {
            err = "a load range needs a positive increment";
            return NR_E_BATCH;
        }
        digits = dec_digits(2 + (int)((cmax - cmin) / cinc) + (int)((smax - smin) / sinc));
        const int max_iter = 20 * 2 * info.n_nodes + 1000;
        for (double comp = cmin; comp < cmax + kFltTol; comp += cinc)
            for (double shear = smin; shear < smax + kFltTol; shear += sinc) {
                comps.push_back(comp);
                shears.push_back(shear);
            }
        u.assign(comps.size(), std::vector<double>(2 * (size_t)info.n_nodes));
        jobs.resize(comps.size());
        reps.resize(comps.size());
        for (size_t k = 0; k < comps.size(); k++) {
            nr_linear_job &jb = jobs[k];
            jb.net = net;
            jb.ks = spec.ks;
            jb.kb = spec.kb;
            jb.mu = spec.mu;
            jb.shear = affine ? shears[k] / yext : shears[k];
            jb.compression = affine ? comps[k] / yext : comps[k];
            jb.affine = affine ? 1 : 0;
            jb.tol = 1e-12;
            jb.max_iter = max_iter;
            jb.u = u[k].data();
        }
        return NR_OK;
    }

    size_t rows() const { return jobs.size() * 2 * (size_t)n_nodes; }

    void write() const
    {
        FILE *lf = std::fopen(jf->log.c_str(), "w");
        FILE *ef = std::fopen(jf->edata.c_str(), "w");
        for (size_t k = 0; k < jobs.size(); k++) {
            const std::string name = report_name(jf->disp, digits, (int)k + 1, "disp");
            FILE *df = std::fopen(name.c_str(), "w");
            if (df) {
                for (int i = 0; i < n_nodes; i++)
                    std::fprintf(df, "%2.10le\t%2.10le\n", u[k][2 * i], u[k][2 * i + 1]);
                std::fclose(df);
            }
            std::fprintf(lf ? lf : stdout, "Residual RMS force: %2.10le\n", reps[k].residual);
            FILE *out = ef ? ef : stdout;
            if (!itemize)
                std::fprintf(out, "%12.8lf\t%12.8lf\t%12.10le\n", comps[k], shears[k],
                             reps[k].e_stretch + reps[k].e_bend);
            else
                std::fprintf(out, "%12.8lf\t%12.8lf\t%12.10le\t%12.10le\n", comps[k], shears[k],
                             reps[k].e_stretch, reps[k].e_bend);
        }
        if (ef) std::fclose(ef);
        if (lf) {
            std::fprintf(lf, "Elapsed time: %d seconds.\n",
                         (int)std::difftime(std::time(nullptr), start));
            std::fclose(lf);
        }
    }
};

}  // namespace nr

// Rows (DOFs x load steps) of one device batch of the linear protocol: a
// directory's networks go to the device in groups of about this many.
constexpr size_t kLinearBatchRows = (size_t)1 << 25;

extern "C" int nr_linear_batch(const char *batch_file, int strain_mode, int affine, int itemize)
{
    return nr::guarded(NR_E_BATCH, [&]() -> int {
        std::string err;
        nr::LinearBatchSpec b;
        if (!batch_file) return NR_E_ARG;
        int rc = nr::parse_linear_batch(batch_file, b, err);
        if (rc) {
            std::fprintf(stderr, "%s\nFile parsing failed.\n", err.c_str());
            return rc;
        }
        std::vector<nr::JobFiles> files;
        rc = nr::file_lists(b.base, b.tag, files, err);
        if (rc) {
            std::fprintf(stderr, "%s\nFile lists could not be read.\n", err.c_str());
            return rc;
        }
        // Every load step of every network is an independent solve: the
        // networks' steps go to the device together, in groups of about
        // kLinearBatchRows rows, and each network's files are written once
        // its group is solved; errors are reported in the file order.
        int first_err = NR_OK;
        auto note = [&](const nr::JobFiles &f, int code, const std::string &msg) {
            std::fprintf(stderr, "%s: %s\n", f.dat.c_str(), msg.empty() ? "failed" : msg.c_str());
            if (!first_err) first_err = code;
        };
        for (size_t k0 = 0; k0 < files.size();) {
            std::vector<std::unique_ptr<nr::LinearRun>> group;
            size_t rows = 0;
            for (; k0 < files.size() && (group.empty() || rows < kLinearBatchRows); k0++) {
                auto run = std::make_unique<nr::LinearRun>();
                std::string e;
                const int prc = run->plan(files[k0], b, strain_mode != 0, affine != 0,
                                          itemize != 0, e);
                if (prc) {
                    note(files[k0], prc, e);
                    continue;
                }
                rows += run->rows();
                group.push_back(std::move(run));
            }
            std::vector<nr_linear_job> jobs;
            for (auto &run : group) jobs.insert(jobs.end(), run->jobs.begin(), run->jobs.end());
            std::vector<nr_linear_report> reps(jobs.size());
            const int src = nr_linear_response_batch((int)jobs.size(), jobs.data(), reps.data());
            const std::string serr = src ? nr_last_error() : "";
            size_t at = 0;
            for (auto &run : group) {
                std::copy(reps.begin() + at, reps.begin() + at + run->jobs.size(), run->reps.begin());
                at += run->jobs.size();
                if (src)
                    note(*run->jf, src, serr);
                else
                    run->write();
            }
        }
        return first_err;
    });
}

extern "C" int nr_relaxation_batch(const char *batch_file, int strain_mode,
                                   const nr_batch_opts *opts)
{
    return nr::guarded(NR_E_BATCH, [&]() -> int {
        std::string err;
        nr::BatchSpec b;
        if (!batch_file) return NR_E_ARG;
        int rc = nr::parse_batch_file(batch_file, b, err);
        if (rc) {
            std::fprintf(stderr, "%s\nParsing of the batch file failed.\n", err.c_str());
            return rc;
        }
        std::vector<nr::JobFiles> jobs;
        rc = nr::file_lists(b.base, b.tag, jobs, err);
        if (rc) {
            std::fprintf(stderr, "%s\nFile lists could not be read.\n", err.c_str());
            return rc;
        }
        int first_err = NR_OK;
        nr::run_protocol_batched(jobs, b, strain_mode != 0, opts, first_err);
        return first_err;
    });
}
