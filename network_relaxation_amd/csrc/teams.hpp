// Team layout of the persistent relax grid.
#pragma once

#include <cmath>
#include <cstdint>
#include <vector>

#include "device.hpp"

namespace nr {

struct TeamLayout {
    int grid = 0;                   // workgroups launched
    int n_teams = 0;
    int n_parts = 0;                // partial-sum slots (sum of team sizes)
    int n_ids = 0;                  // barrier/mailbox ids (split trees, 2g - 1 per team)
    int cls_of_log2g[kMaxLog2Team]; // class of teams of 2^i workgroups, -1 none
    bool cooperative = false;       // some team has more than one member
    double makespan = 0.0;          // predicted, in cost-model units
    std::vector<BlockRole> roles;   // per workgroup
    std::vector<int> class_jobs;    // networks per class, largest first
    std::vector<int> class_off;     // [n_classes + 1]
    std::vector<int> team_of_net;   // team size used for each network
};

// Team size of a network: the largest power of two not above
// ceil(n_tiles / tiles_per_member) and the cap.  It depends only on the
// network and the options, so a network is relaxed by the same arithmetic
// alone and inside any batch.
int team_size(int n_tiles, int cap, int tiles_per_member);

// Cost model of one relax, fitted to per-network timelines of a 512-network
// batch on MI355X: a FIRE step of a network costs the tile passes of its
// busiest member (~4.6 us each) plus the team's wait for its slowest member
// and the partial-sum exchange, ~0.45 tile passes per doubling of the team.
// `steps` is the expected step count.
double job_cost(double steps, int n_tiles, int g);
// Expected FIRE steps of a relax with no history, fitted to 512 relaxed
// triangular networks (L 16..192, bond fraction p 0.55..0.75, shear 0..0.4):
// steps ~ L^0.71 exp(-8.2 p), L = sqrt(N), p = mean coordination / 6
// (R^2 = 0.79 on log steps; L alone explains 0.34).  Sparser networks relax
// far more slowly, which a size-only estimate misses.
inline double expected_steps(int n_nodes, int n_bonds)
{
    const double n = n_nodes > 0 ? (double)n_nodes : 1.0;
    const double p = (2.0 * n_bonds) / (6.0 * n);
    return std::exp(11.70 + 0.707 * std::log(std::sqrt(n)) - 8.24 * p);
}

// One job of a launch plan: its team size, tile count and expected steps.
struct JobSpec {
    int team, n_tiles;
    double steps;
};

// Splits `resident` workgroups between the team-size classes present in the
// jobs: every class gets one team, then teams go one at a time to the class
// whose predicted finish (longest-processing-time-first over its teams) is
// latest, while they fit.  A class never gets more teams than jobs.
TeamLayout plan_teams(const std::vector<JobSpec> &jobs, int resident);

// Like plan_teams, but when the classes cannot all be co-resident (their
// team sizes add up to more than `resident`) the classes are split over
// several launches, largest teams first.  Job lists hold batch indices.
std::vector<TeamLayout> plan_launches(const std::vector<JobSpec> &jobs, int resident);

// Predicted time (cost-model units: plan_launches makespans added over
// launches) of relaxing the networks with this tiles-per-member rule.
double predict_time(const std::vector<int> &n_tiles, const std::vector<double> &steps, int cap,
                    int resident, int tiles_per_member);

// Tiles per member that minimises the predicted time of the job set over
// `candidates` (plan_launches makespans added over launches).
int tune_tiles_per_member(const std::vector<int> &n_tiles, const std::vector<double> &steps,
                          int cap, int resident, const std::vector<int> &candidates,
                          double *predicted);

}  // namespace nr
