// The XML relaxation job format that the reference declares but never parses
// (src/large_strain_types.hpp: element names :36-39, children :42-44,
// attributes :48-50, RelaxScheme/Run/RelaxBatch :99-168, the format's
// description :226-250; parse_relaxation_batch :266-273 is an empty stub and
// the file needs libxml2).  A job:
//
//   <rbatch>
//     <name input="dir/base" tag="t"/>
//     <mech_params stretch="1.0" bend="1e-3"/>
//     <rscheme type="FIRE" pars="0.1 0.99 0.1 1.1 0.5 5 1.0 1e-10 1000000"/>
//     <run>
//       <axial min="0" max="0" inc="0"/>
//       <shear min="0.01" max="0.05" inc="0.01"/>
//       <report edata="true" log="true" disp="false"/>
//     </run>
//     <run> ... </run>
//   </rbatch>
//
// Element and attribute names follow the reference's tables; `run` is a child
// of `rbatch` as the description says (the children table omits it, which
// would leave no way to give a run).  Meanings, where the reference leaves
// them open, are those of the 6-line batch file: `input`/`tag` line 1,
// stretch/bend line 2, FIRE `pars` lines 3 and 4 (alpha0 f_alpha dt_max f_inc
// f_dec N_min mass fcut max_steps), `ConjGrad` pars "fcut max_iters" (or the
// nine FIRE numbers, of which it uses fcut and max_steps), `axial` line 5
// (compression), `shear` line 6.  Runs execute in order on each network's
// resident state; with one run the outputs are the batch protocol's, with
// several, run k's names carry "_run<k>" before "_log.txt" / "_edata.txt" /
// the .pos counter.  A report flag of false skips that file.
#include <cctype>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "../../include/nrelax.h"
#include "internal.hpp"
#include "protocol.hpp"

namespace nr {
namespace {

struct XNode {
    std::string name;
    std::vector<std::pair<std::string, std::string>> attrs;
    std::vector<XNode> kids;
    int line = 0;
};

// A small reader for the job files: elements with quoted attributes, the
// declaration, comments and DOCTYPE skipped, no text content.
class XmlReader {
public:
    explicit XmlReader(const std::string &t) : s_(t) {}

    bool parse(XNode &root, std::string &err)
    {
        if (!skip_misc(err)) return false;
        if (p_ >= s_.size() || s_[p_] != '<') return fail_at("expected the root element", err);
        if (!element(root, 0, err)) return false;
        if (!skip_misc(err)) return false;
        if (p_ != s_.size()) return fail_at("text after the root element", err);
        return true;
    }

private:
    const std::string &s_;
    size_t p_ = 0;
    int line_ = 1;

    bool fail_at(const std::string &what, std::string &err)
    {
        err = "line " + std::to_string(line_) + ": " + what;
        return false;
    }
    void adv(size_t n)
    {
        for (size_t k = 0; k < n && p_ < s_.size(); k++)
            if (s_[p_++] == '\n') line_++;
    }
    bool starts(const char *t) const { return s_.compare(p_, std::strlen(t), t) == 0; }
    void ws()
    {
        while (p_ < s_.size() && std::isspace((unsigned char)s_[p_])) adv(1);
    }
    bool skip_to(const char *end, std::string &err)
    {
        size_t q = s_.find(end, p_);
        if (q == std::string::npos) return fail_at(std::string("unterminated, missing ") + end, err);
        adv(q + std::strlen(end) - p_);
        return true;
    }
    // whitespace, <?...?>, <!--...-->, <!DOCTYPE ...>
    bool skip_misc(std::string &err)
    {
        for (;;) {
            ws();
            if (starts("\xEF\xBB\xBF")) adv(3);
            else if (starts("<?")) { if (!skip_to("?>", err)) return false; }
            else if (starts("<!--")) { if (!skip_to("-->", err)) return false; }
            else if (starts("<!")) { if (!skip_to(">", err)) return false; }
            else return true;
        }
    }
    std::string name()
    {
        size_t q = p_;
        while (q < s_.size() && (std::isalnum((unsigned char)s_[q]) || s_[q] == '_' ||
                                 s_[q] == '-' || s_[q] == ':' || s_[q] == '.'))
            q++;
        std::string n = s_.substr(p_, q - p_);
        adv(q - p_);
        return n;
    }
    static bool unescape(const std::string &v, std::string &out)
    {
        static const char *ent[][2] = {{"&lt;", "<"}, {"&gt;", ">"}, {"&amp;", "&"},
                                       {"&quot;", "\""}, {"&apos;", "'"}};
        out.clear();
        for (size_t i = 0; i < v.size();) {
            if (v[i] != '&') {
                out.push_back(v[i++]);
                continue;
            }
            bool ok = false;
            for (auto &e : ent)
                if (v.compare(i, std::strlen(e[0]), e[0]) == 0) {
                    out += e[1];
                    i += std::strlen(e[0]);
                    ok = true;
                    break;
                }
            if (!ok) return false;
        }
        return true;
    }
    bool element(XNode &n, int depth, std::string &err)
    {
        if (depth > 16) return fail_at("elements nested too deeply", err);
        n.line = line_;
        adv(1);  // '<'
        n.name = name();
        if (n.name.empty()) return fail_at("an element without a name", err);
        for (;;) {
            ws();
            if (starts("/>")) {
                adv(2);
                return true;
            }
            if (starts(">")) {
                adv(1);
                break;
            }
            std::string a = name();
            if (a.empty()) return fail_at("malformed attribute in <" + n.name + ">", err);
            ws();
            if (!starts("=")) return fail_at("attribute " + a + " without a value", err);
            adv(1);
            ws();
            if (p_ >= s_.size() || (s_[p_] != '"' && s_[p_] != '\''))
                return fail_at("attribute " + a + " value is not quoted", err);
            const char q = s_[p_];
            adv(1);
            size_t e = s_.find(q, p_);
            if (e == std::string::npos) return fail_at("unterminated attribute " + a, err);
            std::string v;
            if (!unescape(s_.substr(p_, e - p_), v))
                return fail_at("unknown entity in attribute " + a, err);
            adv(e + 1 - p_);
            for (auto &kv : n.attrs)
                if (kv.first == a) return fail_at("attribute " + a + " given twice", err);
            n.attrs.push_back({a, v});
        }
        for (;;) {  // content: child elements only
            if (!skip_misc(err)) return false;
            if (p_ >= s_.size()) return fail_at("<" + n.name + "> is not closed", err);
            if (starts("</")) {
                adv(2);
                std::string c = name();
                ws();
                if (c != n.name || !starts(">"))
                    return fail_at("</" + c + "> does not close <" + n.name + ">", err);
                adv(1);
                return true;
            }
            if (s_[p_] != '<') return fail_at("text inside <" + n.name + ">", err);
            n.kids.emplace_back();
            if (!element(n.kids.back(), depth + 1, err)) return false;
        }
    }
};

struct XmlRun {
    double cmin = 0, cmax = 0, cinc = 0, smin = 0, smax = 0, sinc = 0;
    unsigned report = kReportAll;
};

struct XmlJob {
    BatchSpec spec;  // base, tag, ks, kb, FIRE/CG parameters (ramps per run)
    std::vector<XmlRun> runs;
};

// large_strain_types.hpp :36-50, with run under rbatch (:245-249).
struct ElemRule {
    const char *name;
    std::vector<const char *> kids, attrs;
};
const std::vector<ElemRule> &rules()
{
    static const std::vector<ElemRule> r = {
        {"rbatch", {"name", "mech_params", "rscheme", "run"}, {}},
        {"name", {}, {"input", "tag"}},
        {"mech_params", {}, {"stretch", "bend"}},
        {"rscheme", {}, {"type", "pars"}},
        {"run", {"axial", "shear", "report"}, {}},
        {"axial", {}, {"min", "max", "inc"}},
        {"shear", {}, {"min", "max", "inc"}},
        {"report", {}, {"edata", "log", "disp"}},
    };
    return r;
}

const ElemRule *rule(const std::string &n)
{
    for (const ElemRule &r : rules())
        if (n == r.name) return &r;
    return nullptr;
}

bool check_tree(const XNode &n, std::string &err)
{
    const ElemRule *r = rule(n.name);
    const std::string at = "line " + std::to_string(n.line) + ": ";
    if (!r) {
        err = at + "unknown element <" + n.name + ">";
        return false;
    }
    for (auto &kv : n.attrs) {
        bool ok = false;
        for (const char *a : r->attrs) ok = ok || kv.first == a;
        if (!ok) {
            err = at + "<" + n.name + "> has no attribute " + kv.first;
            return false;
        }
    }
    for (const XNode &k : n.kids) {
        bool ok = false;
        for (const char *c : r->kids) ok = ok || k.name == c;
        if (!ok) {
            err = "line " + std::to_string(k.line) + ": <" + k.name + "> cannot appear in <" +
                  n.name + ">";
            return false;
        }
        if (!check_tree(k, err)) return false;
    }
    return true;
}

const std::string *attr(const XNode &n, const char *a)
{
    for (auto &kv : n.attrs)
        if (kv.first == a) return &kv.second;
    return nullptr;
}

bool number(const XNode &n, const char *a, double &v, bool required, std::string &err)
{
    const std::string *s = attr(n, a);
    if (!s) {
        if (!required) return true;
        err = "line " + std::to_string(n.line) + ": <" + n.name + "> needs " + a;
        return false;
    }
    char *end = nullptr;
    v = std::strtod(s->c_str(), &end);
    while (end && *end && std::isspace((unsigned char)*end)) end++;
    if (s->empty() || !end || *end || !std::isfinite(v)) {
        err = "line " + std::to_string(n.line) + ": " + a + "=\"" + *s + "\" is not a finite number";
        return false;
    }
    return true;
}

bool flag(const XNode &n, const char *a, bool &v, std::string &err)
{
    const std::string *s = attr(n, a);
    if (!s) return true;  // absent: reported
    if (*s == "true" || *s == "1" || *s == "yes") v = true;
    else if (*s == "false" || *s == "0" || *s == "no") v = false;
    else {
        err = "line " + std::to_string(n.line) + ": " + a + "=\"" + *s + "\" is not true/false";
        return false;
    }
    return true;
}

bool ramp(const XNode &n, double &mn, double &mx, double &inc, std::string &err)
{
    if (!number(n, "min", mn, true, err) || !number(n, "max", mx, true, err) ||
        !number(n, "inc", inc, true, err))
        return false;
    if (!valid_range(mn, mx, inc)) {  // as batch lines 5 and 6
        err = "line " + std::to_string(n.line) + ": invalid <" + n.name + "> range";
        return false;
    }
    return true;
}

int load_job(const std::string &path, XmlJob &job, std::string &err)
{
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) {
        err = "The file " + path + " could not be read.";
        return NR_E_OPEN;
    }
    std::ostringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    XNode root;
    if (!XmlReader(text).parse(root, err) || !check_tree(root, err)) return NR_E_BATCH;
    if (root.name != "rbatch") {
        err = "the root element must be <rbatch>";
        return NR_E_BATCH;
    }
    const XNode *nm = nullptr, *mp = nullptr, *rs = nullptr;
    for (const XNode &k : root.kids) {
        const XNode **slot = k.name == "name" ? &nm : k.name == "mech_params" ? &mp
                             : k.name == "rscheme" ? &rs : nullptr;
        if (!slot) continue;
        if (*slot) {
            err = "line " + std::to_string(k.line) + ": <" + k.name + "> given twice";
            return NR_E_BATCH;
        }
        *slot = &k;
    }
    if (!nm || !mp || !rs) {
        err = "<rbatch> needs <name>, <mech_params> and <rscheme>";
        return NR_E_BATCH;
    }
    BatchSpec &b = job.spec;
    b = BatchSpec();
    const std::string *input = attr(*nm, "input");
    if (!input || input->empty()) {
        err = "<name> needs input";
        return NR_E_BATCH;
    }
    b.base = *input;
    if (const std::string *t = attr(*nm, "tag")) b.tag = *t;
    if (!number(*mp, "stretch", b.ks, true, err) || !number(*mp, "bend", b.kb, true, err))
        return NR_E_BATCH;
    if (b.ks < 0 || b.kb < 0) {
        err = "stretch and bend must be >= 0";
        return NR_E_BATCH;
    }
    const std::string *type = attr(*rs, "type"), *pars = attr(*rs, "pars");
    if (!type || !pars) {
        err = "<rscheme> needs type and pars";
        return NR_E_BATCH;
    }
    std::vector<double> pv;
    {
        std::istringstream ps(*pars);
        std::string tok;
        while (ps >> tok) {
            char *end = nullptr;
            const double v = std::strtod(tok.c_str(), &end);
            if (!end || *end || !std::isfinite(v)) {
                err = "pars: \"" + tok + "\" is not a finite number";
                return NR_E_BATCH;
            }
            pv.push_back(v);
        }
    }
    if (*type == "FIRE" || (*type == "ConjGrad" && pv.size() >= 9)) {
        if (pv.size() < 9) {
            err = "FIRE pars need 9 numbers: alpha0 f_alpha dt_max f_inc f_dec N_min mass "
                  "fcut max_steps";
            return NR_E_BATCH;
        }
        if (!(pv[5] >= 0 && pv[5] <= INT_MAX) || !(pv[8] >= 0 && pv[8] <= INT_MAX)) {
            err = "N_min and max_steps must lie in [0, 2^31 - 1]";
            return NR_E_BATCH;
        }
        for (int k = 0; k < 5; k++) b.fp[k] = pv[k];
        b.nmin = (int)pv[5];
        b.mass = pv[6];
        b.fcut = pv[7];
        b.nmax = (int)pv[8];
    } else if (*type == "ConjGrad") {
        if (pv.size() != 2) {
            err = "ConjGrad pars are \"fcut max_iters\" (or the 9 FIRE numbers)";
            return NR_E_BATCH;
        }
        if (!(pv[1] >= 0 && pv[1] <= INT_MAX)) {
            err = "max_iters must lie in [0, 2^31 - 1]";
            return NR_E_BATCH;
        }
        const double fire[5] = {0.1, 0.99, 0.1, 1.1, 0.5};
        for (int k = 0; k < 5; k++) b.fp[k] = fire[k];
        b.nmin = 5;
        b.mass = 1.0;
        b.fcut = pv[0];
        b.nmax = (int)pv[1];
    } else {
        err = "rscheme type must be FIRE or ConjGrad (RelaxMethods, large_strain_types.hpp:34)";
        return NR_E_BATCH;
    }
    b.method = *type == "ConjGrad" ? NR_METHOD_CG : NR_METHOD_FIRE;
    if (!(b.mass > 0) || b.nmax < 0 || b.fcut < 0) {
        err = "mass must be > 0, fcut and max_steps >= 0";
        return NR_E_BATCH;
    }
    job.runs.clear();
    for (const XNode &k : root.kids) {
        if (k.name != "run") continue;
        XmlRun r;
        bool seen[3] = {false, false, false};
        for (const XNode &c : k.kids) {
            const int w = c.name == "axial" ? 0 : c.name == "shear" ? 1 : 2;
            if (seen[w]) {
                err = "line " + std::to_string(c.line) + ": <" + c.name + "> given twice in <run>";
                return NR_E_BATCH;
            }
            seen[w] = true;
            if (w == 0 && !ramp(c, r.cmin, r.cmax, r.cinc, err)) return NR_E_BATCH;
            if (w == 1 && !ramp(c, r.smin, r.smax, r.sinc, err)) return NR_E_BATCH;
            if (w == 2) {
                bool e = true, l = true, d = true;
                if (!flag(c, "edata", e, err) || !flag(c, "log", l, err) || !flag(c, "disp", d, err))
                    return NR_E_BATCH;
                r.report = (e ? kReportEdata : 0u) | (l ? kReportLog : 0u) | (d ? kReportDisp : 0u);
            }
        }
        job.runs.push_back(r);
    }
    if (job.runs.empty()) {
        err = "<rbatch> needs at least one <run>";
        return NR_E_BATCH;
    }
    return NR_OK;
}

// Output names of run k (1-based) of `count`.
JobFiles run_files(const JobFiles &jf, int k, int count)
{
    if (count == 1) return jf;
    JobFiles r = jf;
    const std::string s = "_run" + std::to_string(k);
    auto insert = [&](std::string &name, const char *suffix) {
        const size_t n = std::strlen(suffix);
        if (name.size() >= n) name.insert(name.size() - n, s);
    };
    insert(r.log, "_log.txt");
    insert(r.edata, "_edata.txt");
    r.disp += s;
    return r;
}

std::string plan_text(const XmlJob &job, const std::vector<JobFiles> &jobs)
{
    const BatchSpec &b = job.spec;
    std::ostringstream o;
    o.precision(17);
    o << "input " << b.base << " tag " << (b.tag.empty() ? "-" : b.tag) << "\n";
    o << "stretch " << b.ks << " bend " << b.kb << "\n";
    o << "method " << (b.method == NR_METHOD_CG ? "ConjGrad" : "FIRE") << " pars";
    for (double v : b.fp) o << " " << v;
    o << " " << b.nmin << " " << b.mass << " " << b.fcut << " " << b.nmax << "\n";
    for (size_t k = 0; k < job.runs.size(); k++) {
        const XmlRun &r = job.runs[k];
        o << "run " << k + 1 << " axial " << r.cmin << " " << r.cmax << " " << r.cinc << " shear "
          << r.smin << " " << r.smax << " " << r.sinc << " report"
          << ((r.report & kReportEdata) ? " edata" : "") << ((r.report & kReportLog) ? " log" : "")
          << ((r.report & kReportDisp) ? " disp" : "") << "\n";
    }
    for (const JobFiles &j : jobs)
        for (size_t k = 0; k < job.runs.size(); k++) {
            const JobFiles f = run_files(j, (int)k + 1, (int)job.runs.size());
            o << "network " << j.dat << " run " << k + 1 << " log " << f.log << " edata " << f.edata
              << " disp " << f.disp << "\n";
        }
    return o.str();
}

}  // namespace
}  // namespace nr

extern "C" int nr_xml_job_plan(const char *job_file, char *plan, int cap)
{
    return nr::guarded(NR_E_BATCH, [&]() -> int {
        if (!job_file) return nr::fail(NR_E_ARG, "null argument");
        nr::XmlJob job;
        std::string err;
        int rc = nr::load_job(job_file, job, err);
        if (rc) return nr::fail(rc, err);
        std::vector<nr::JobFiles> jobs;
        rc = nr::file_lists(job.spec.base, job.spec.tag, jobs, err);
        if (rc) return nr::fail(rc, err);
        const std::string text = nr::plan_text(job, jobs);
        if (plan && cap > 0) std::snprintf(plan, (size_t)cap, "%s", text.c_str());
        return (int)text.size() < cap || !plan ? NR_OK : nr::fail(NR_E_ARG, "plan buffer too small");
    });
}

extern "C" int nr_xml_batch(const char *job_file, int strain_mode, const nr_batch_opts *opts)
{
    return nr::guarded(NR_E_BATCH, [&]() -> int {
        if (!job_file) return NR_E_ARG;
        nr::XmlJob job;
        std::string err;
        int rc = nr::load_job(job_file, job, err);
        if (rc) {
            std::fprintf(stderr, "%s\nParsing of the job file failed.\n", err.c_str());
            return nr::fail(rc, err);
        }
        std::vector<nr::JobFiles> jobs;
        rc = nr::file_lists(job.spec.base, job.spec.tag, jobs, err);
        if (rc) {
            std::fprintf(stderr, "%s\nFile lists could not be read.\n", err.c_str());
            return nr::fail(rc, err);
        }
        int first_err = NR_OK;
        for (const nr::JobFiles &j : jobs) {
            nr::ProtoNet pn;
            rc = nr::proto_open(j, job.spec, opts, pn, err);
            for (size_t k = 0; rc == NR_OK && k < job.runs.size(); k++) {
                nr::BatchSpec b = job.spec;
                const nr::XmlRun &r = job.runs[k];
                b.cmin = r.cmin;
                b.cmax = r.cmax;
                b.cinc = r.cinc;
                b.smin = r.smin;
                b.smax = r.smax;
                b.sinc = r.sinc;
                rc = nr::proto_ramps(pn, nr::run_files(j, (int)k + 1, (int)job.runs.size()), b,
                                     strain_mode != 0, r.report, err);
            }
            if (rc) {
                std::fprintf(stderr, "%s: %s\n", j.dat.c_str(), err.c_str());
                if (!first_err) first_err = rc;
            }
        }
        return first_err;
    });
}
