// Team layout of the persistent relax grid: size classes, workgroup shares,
// per-workgroup roles and the class job lists.
#include "teams.hpp"

#include <algorithm>
#include <cmath>
#include <functional>
#include <map>

namespace nr {

int team_size(int n_tiles, int cap, int tiles_per_member)
{
    const int r = std::max(1, tiles_per_member);
    int lim = std::max(1, std::min((n_tiles + r - 1) / r, cap));
    int g = 1;
    while (2 * g <= lim) g *= 2;
    return g;
}

TeamLayout plan_teams(const std::vector<int> &team, const std::vector<int> &n_nodes,
                      int resident)
{
    TeamLayout T;
    const int count = (int)team.size();
    T.team_of_net = team;
    std::map<int, std::vector<int>> classes;  // team size -> networks
    for (int k = 0; k < count; k++) classes[team[k]].push_back(k);
    std::vector<int> gs;
    std::vector<double> work;
    std::vector<int> cnt;
    double total = 0.0;
    for (auto &kv : classes) {
        double w = 0.0;
        for (int k : kv.second) w += std::pow((double)n_nodes[k], 1.5);
        gs.push_back(kv.first);
        work.push_back(w);
        cnt.push_back((int)kv.second.size());
        total += w;
    }
    const int nc = (int)gs.size();
    // teams per class: at least one, at most one per network, blocks ~ work
    std::vector<int> teams(nc, 1);
    int used = 0;
    for (int c = 0; c < nc; c++) used += gs[c];
    while (true) {
        int best = -1;
        double best_deficit = 0.0;
        for (int c = 0; c < nc; c++) {
            if (teams[c] >= cnt[c] || used + gs[c] > resident) continue;
            double share = work[c] / total * resident;
            double deficit = share - teams[c] * gs[c];
            if (deficit > best_deficit) {
                best_deficit = deficit;
                best = c;
            }
        }
        if (best < 0) break;
        teams[best]++;
        used += gs[best];
    }
    // roles, job lists
    int team_id = 0, part = 0;
    T.class_off.push_back(0);
    int c = 0;
    for (auto &kv : classes) {
        std::vector<int> jobs = kv.second;
        std::stable_sort(jobs.begin(), jobs.end(),
                         [&](int a, int b) { return n_nodes[a] > n_nodes[b]; });
        T.class_jobs.insert(T.class_jobs.end(), jobs.begin(), jobs.end());
        T.class_off.push_back((int)T.class_jobs.size());
        for (int t = 0; t < teams[c]; t++) {
            for (int m = 0; m < gs[c]; m++) {
                BlockRole r = {};
                r.team = team_id;
                r.member = m;
                r.g = gs[c];
                r.cls = c;
                r.part = part;
                T.roles.push_back(r);
            }
            team_id++;
            part += gs[c];
            if (gs[c] > 1) T.cooperative = true;
        }
        c++;
    }
    T.grid = (int)T.roles.size();
    T.n_teams = team_id;
    T.n_parts = part;
    return T;
}

std::vector<TeamLayout> plan_launches(const std::vector<int> &team,
                                      const std::vector<int> &n_nodes, int resident)
{
    const int count = (int)team.size();
    std::map<int, std::vector<int>, std::greater<int>> classes;
    for (int k = 0; k < count; k++) classes[team[k]].push_back(k);
    // first-fit of classes (largest team first) into launches
    std::vector<std::vector<int>> groups;
    std::vector<int> room;
    for (auto &kv : classes) {
        int placed = -1;
        for (size_t i = 0; i < groups.size(); i++)
            if (room[i] >= kv.first) {
                placed = (int)i;
                break;
            }
        if (placed < 0) {
            groups.push_back({});
            room.push_back(resident);
            placed = (int)groups.size() - 1;
        }
        room[placed] -= kv.first;
        groups[placed].insert(groups[placed].end(), kv.second.begin(), kv.second.end());
    }
    std::vector<TeamLayout> out;
    for (auto &grp : groups) {
        std::sort(grp.begin(), grp.end());
        std::vector<int> t, n;
        for (int k : grp) {
            t.push_back(team[k]);
            n.push_back(n_nodes[k]);
        }
        TeamLayout L = plan_teams(t, n, resident);
        for (int &j : L.class_jobs) j = grp[j];
        std::vector<int> tm(count, 0);
        for (size_t i = 0; i < grp.size(); i++) tm[grp[i]] = L.team_of_net[i];
        L.team_of_net = tm;
        out.push_back(std::move(L));
    }
    return out;
}

}  // namespace nr
