// Team layout of the persistent relax grid.
#pragma once

#include <cstdint>
#include <vector>

#include "device.hpp"

namespace nr {

struct TeamLayout {
    int grid = 0;                   // workgroups launched
    int n_teams = 0;
    int n_parts = 0;                // partial-sum slots (sum of team sizes)
    bool cooperative = false;       // some team has more than one member
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

// Splits `resident` workgroups between the team-size classes (team[k] is
// network k's team size) present in the
// batch in proportion to their estimated work (nodes^1.5 per network), each
// class getting at least one team and at most one team per network.
TeamLayout plan_teams(const std::vector<int> &team, const std::vector<int> &n_nodes,
                      int resident);

// Like plan_teams, but when the classes cannot all be co-resident (their
// team sizes add up to more than `resident`) the classes are split over
// several launches, largest teams first.  Job lists hold batch indices.
std::vector<TeamLayout> plan_launches(const std::vector<int> &team,
                                      const std::vector<int> &n_nodes, int resident);

}  // namespace nr
