// Batch-file protocol of the reference's `lsrelax`.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/nrelax.h"

namespace nr {

struct BatchSpec {
    std::string base, tag;
    double ks = 0, kb = 0;
    double fp[5] = {0, 0, 0, 0, 0};
    int nmin = 0;
    double mass = 0, fcut = 0;
    int nmax = 0;
    double cmin = 0, cmax = 0, cinc = 0;
    double smin = 0, smax = 0, sinc = 0;
    int criterion = NR_CRIT_RMS;
    int method = NR_METHOD_FIRE;
    double kappa_lo = 0, kappa_hi = 0;
    uint64_t kappa_seed = 0;
};

struct JobFiles {
    std::string dat, log, edata, disp;
};

// Most increments a strain ramp (min max inc) may take.
constexpr double kMaxRampSteps = 1e7;
// A ramp's numbers are finite, 0 <= min <= max, inc >= 0, and the ramp takes
// at most kMaxRampSteps increments.
bool valid_range(double mn, double mx, double inc);
int parse_batch_file(const std::string &path, BatchSpec &b, std::string &err);
// ext: the input extension ("dat" networks, "nmc" network + gel composites)
int file_lists(const std::string &input, const std::string &tag, std::vector<JobFiles> &jobs,
               std::string &err, const std::string &ext = "dat");
int dec_digits(int value);
std::string report_name(const std::string &base, int digits, int count, const std::string &ext);
int run_protocol(const JobFiles &jf, const BatchSpec &b, bool strain_mode,
                 const nr_batch_opts *opts, std::string &err);
// relaxation_batch: every network of jobs in one device batch (groups bounded
// by node count); first_err receives the first failure's code.
int run_protocol_batched(const std::vector<JobFiles> &jobs, const BatchSpec &b, bool strain_mode,
                         const nr_batch_opts *opts, int &first_err);

// Outputs of a ramp (the XML job format's report flags).
constexpr unsigned kReportEdata = 1, kReportLog = 2, kReportDisp = 4, kReportAll = 7;
// One network of a batch, resident on the device for all its ramps.
struct ProtoNet {
    nr_network *n = nullptr;
    nr_batch *b = nullptr;
    nr_network_info info{};
    ProtoNet() = default;
    ProtoNet(const ProtoNet &) = delete;
    ProtoNet &operator=(const ProtoNet &) = delete;
    ~ProtoNet()
    {
        nr_batch_destroy(b);
        nr_network_free(n);
    }
};
int proto_open(const JobFiles &jf, const BatchSpec &b, const nr_batch_opts *opts, ProtoNet &pn,
               std::string &err);
int proto_ramps(ProtoNet &pn, const JobFiles &jf, const BatchSpec &b, bool strain_mode,
                unsigned report, std::string &err);

// Batch file of the reference's small-strain program `sp_relax`
// (sparse_relax.cpp:920-1035): base [tag]; ks kb mu; compression min max
// inc; shear min max inc.
struct LinearBatchSpec {
    std::string base, tag;
    double ks = 0, kb = 0, mu = 0;
    double cmin = 0, cmax = 0, cinc = 0;
    double smin = 0, smax = 0, sinc = 0;
};
int parse_linear_batch(const std::string &path, LinearBatchSpec &b, std::string &err);

}  // namespace nr
