// Team layout of the persistent relax grid: size classes, workgroup shares,
// per-workgroup roles and the class job lists.
#include "teams.hpp"

#include <algorithm>
#include <functional>
#include <limits>
#include <map>
#include <queue>

namespace nr {

int team_size(int n_tiles, int cap, int tiles_per_member)
{
    const int r = std::max(1, tiles_per_member);
    int lim = std::max(1, std::min((n_tiles + r - 1) / r, cap));
    int g = 1;
    while (2 * g <= lim) g *= 2;
    return g;
}

double job_cost(double steps, int n_tiles, int g)
{
    const double passes = (double)((n_tiles + g - 1) / g);
    return steps * (passes + 0.45 * std::log2((double)g));
}

namespace {

// Finish time of `teams` identical teams serving `costs` (sorted descending)
// longest first, each job going to the team that frees up first -- the order
// in which the device queue hands them out.
double lpt_finish(const std::vector<double> &costs, int teams)
{
    std::priority_queue<double, std::vector<double>, std::greater<double>> free_at;
    for (int t = 0; t < teams; t++) free_at.push(0.0);
    double end = 0.0;
    for (double c : costs) {
        double t0 = free_at.top();
        free_at.pop();
        end = std::max(end, t0 + c);
        free_at.push(t0 + c);
    }
    return end;
}

}  // namespace

TeamLayout plan_teams(const std::vector<JobSpec> &jobs, int resident)
{
    TeamLayout T;
    const int count = (int)jobs.size();
    T.team_of_net.resize(count);
    std::map<int, std::vector<int>> classes;  // team size -> jobs
    for (int k = 0; k < count; k++) {
        classes[jobs[k].team].push_back(k);
        T.team_of_net[k] = jobs[k].team;
    }
    const int nc = (int)classes.size();
    std::vector<int> gs, teams(nc, 1);
    std::vector<std::vector<int>> members;
    std::vector<std::vector<double>> costs(nc);
    std::vector<double> finish(nc);
    int used = 0, c = 0;
    for (auto &kv : classes) {
        std::vector<int> js = kv.second;
        const int g = kv.first;
        // longest expected relax first: the queue order on the device
        std::stable_sort(js.begin(), js.end(), [&](int a, int b) {
            return job_cost(jobs[a].steps, jobs[a].n_tiles, g) >
                   job_cost(jobs[b].steps, jobs[b].n_tiles, g);
        });
        for (int k : js) costs[c].push_back(job_cost(jobs[k].steps, jobs[k].n_tiles, g));
        finish[c] = lpt_finish(costs[c], 1);
        gs.push_back(g);
        members.push_back(js);
        used += g;
        c++;
    }
    // one more team at a time to the latest-finishing class that can use it
    while (true) {
        int best = -1;
        for (int i = 0; i < nc; i++) {
            if (teams[i] >= (int)costs[i].size() || used + gs[i] > resident) continue;
            if (best < 0 || finish[i] > finish[best]) best = i;
        }
        if (best < 0) break;
        teams[best]++;
        used += gs[best];
        finish[best] = lpt_finish(costs[best], teams[best]);
    }
    T.makespan = 0.0;
    for (int i = 0; i < nc; i++) T.makespan = std::max(T.makespan, finish[i]);
    // roles, job lists; a team whose class runs out of jobs halves and serves
    // the class of half its size (kernel), so each team owns the ids of its
    // split tree
    int team_id = 0, part = 0, ids = 0;
    for (int i = 0; i < kMaxLog2Team; i++) T.cls_of_log2g[i] = -1;
    T.class_off.push_back(0);
    for (int i = 0; i < nc; i++) {
        T.class_jobs.insert(T.class_jobs.end(), members[i].begin(), members[i].end());
        T.class_off.push_back((int)T.class_jobs.size());
        int lg = 0;
        while ((1 << lg) < gs[i]) lg++;
        if (lg < kMaxLog2Team) T.cls_of_log2g[lg] = i;
        for (int t = 0; t < teams[i]; t++) {
            for (int m = 0; m < gs[i]; m++) {
                BlockRole r = {};
                r.team = team_id;
                r.member = m;
                r.g = gs[i];
                r.cls = i;
                r.part = part;
                r.tree = ids;
                T.roles.push_back(r);
            }
            team_id++;
            part += gs[i];
            ids += 2 * gs[i] - 1;
            if (gs[i] > 1) T.cooperative = true;
        }
    }
    T.n_ids = ids;
    T.grid = (int)T.roles.size();
    T.n_teams = team_id;
    T.n_parts = part;
    return T;
}

std::vector<TeamLayout> plan_launches(const std::vector<JobSpec> &jobs, int resident)
{
    const int count = (int)jobs.size();
    std::map<int, std::vector<int>, std::greater<int>> classes;
    for (int k = 0; k < count; k++) classes[jobs[k].team].push_back(k);
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
        std::vector<JobSpec> sub;
        for (int k : grp) sub.push_back(jobs[k]);
        TeamLayout L = plan_teams(sub, resident);
        for (int &j : L.class_jobs) j = grp[j];
        std::vector<int> tm(count, 0);
        for (size_t i = 0; i < grp.size(); i++) tm[grp[i]] = L.team_of_net[i];
        L.team_of_net = tm;
        out.push_back(std::move(L));
    }
    return out;
}

double predict_time(const std::vector<int> &n_tiles, const std::vector<double> &steps, int cap,
                    int resident, int tiles_per_member)
{
    std::vector<JobSpec> jobs(n_tiles.size());
    for (size_t k = 0; k < n_tiles.size(); k++)
        jobs[k] = {team_size(n_tiles[k], cap, tiles_per_member), n_tiles[k], steps[k]};
    double t = 0.0;
    for (const TeamLayout &L : plan_launches(jobs, resident)) t += L.makespan;
    return t;
}

int tune_tiles_per_member(const std::vector<int> &n_tiles, const std::vector<double> &steps,
                          int cap, int resident, const std::vector<int> &candidates,
                          double *predicted)
{
    int best = candidates.empty() ? 1 : candidates[0];
    double best_t = std::numeric_limits<double>::infinity();
    for (int r : candidates) {
        const double t = predict_time(n_tiles, steps, cap, resident, r);
        if (t < best_t) {
            best_t = t;
            best = r;
        }
    }
    if (predicted) *predicted = best_t;
    return best;
}

}  // namespace nr
