// Admission and batching policy of the request server (serve.hip).  Host
// code only, no HIP: the same header is compiled into a CPU unit test.
//
// The server keeps a FIFO queue of shot-solve requests and a fixed set of
// lanes (one device context and stream each).  Whenever a lane is idle the
// policy picks the next batch:
//   * the oldest request first; later requests of the same batching class
//     (same grid, wave type, mode and acquisition, so they can share one
//     launch) join it in arrival order while the batch has shot slots left;
//   * the batch must fit the device-memory budget: resident models plus what
//     every lane holds (busy lanes their reservation, idle lanes the buffers
//     they keep warm) plus what the chosen lane will hold for the batch;
//     idle lanes may be released to make room, and the batch is shrunk from
//     its tail before it is given up;
//   * if the oldest request does not fit, younger ones may start ahead of it
//     (backfill), but a request is overtaken at most max_bypass times; after
//     that the queue waits for memory to free up.
#pragma once

#include <algorithm>
#include <cstddef>
#include <deque>
#include <functional>
#include <string>
#include <vector>

namespace inv {
namespace serve {

struct Pending {
    long long id = 0;
    std::string cls;   // batching class
    int shots = 1;
    int nt = 1;
    int bypassed = 0;  // times a younger request was admitted ahead of it
};

enum LaneState { LANE_IDLE = 0, LANE_BUSY = 1 };

struct LaneView {
    int state = LANE_IDLE;
    std::string cls;   // class of the context the lane holds ("" = none)
    size_t held = 0;   // device bytes held (idle) or reserved (busy)
};

struct Policy {
    size_t budget = 0;    // device bytes for lanes plus resident models
    int max_slots = 16;   // shots per batch
    int max_bypass = 4;   // times the oldest request may be overtaken
    bool batching = true; // false: one request per batch
};

struct Plan {
    int lane = -1;
    std::vector<size_t> members;  // queue indices, ascending (arrival order)
    int shots = 0, nt = 0;
    size_t need = 0;              // bytes the lane holds while running it
    std::vector<int> trim;        // idle lanes to release first
};

// Bytes a lane context needs for `shots` shots of class `cls` run for nt steps.
using NeedFn = std::function<size_t(const std::string &cls, int shots, int nt)>;

// The lane a batch of class cls should use: an idle lane already holding that
// class (the one holding most), else an empty idle lane, else the idle lane
// holding least.  -1 if no lane is idle.
inline int pick_lane(const std::vector<LaneView> &lanes, const std::string &cls) {
    int best = -1;
    for (int l = 0; l < (int)lanes.size(); l++) {
        if (lanes[l].state != LANE_IDLE || lanes[l].cls != cls || cls.empty()) continue;
        if (best < 0 || lanes[l].held > lanes[best].held) best = l;
    }
    if (best >= 0) return best;
    for (int l = 0; l < (int)lanes.size(); l++) {
        if (lanes[l].state != LANE_IDLE) continue;
        if (best < 0 || lanes[l].held < lanes[best].held) best = l;
    }
    return best;
}

// Try to fit a batch (shots, nt of class cls) on lane `lane`; on success fill
// need and trim.
inline bool fit(const std::vector<LaneView> &lanes, int lane, const std::string &cls, int shots,
                int nt, size_t models_bytes, const Policy &pol, const NeedFn &need_fn,
                size_t &need, std::vector<int> &trim) {
    need = need_fn(cls, shots, nt);
    const LaneView &L = lanes[lane];
    const size_t mine = (L.cls == cls) ? std::max(L.held, need) : need;
    size_t others = models_bytes;
    std::vector<std::pair<size_t, int>> idle;
    for (int l = 0; l < (int)lanes.size(); l++) {
        if (l == lane) continue;
        others += lanes[l].held;
        if (lanes[l].state == LANE_IDLE && lanes[l].held > 0) idle.push_back({lanes[l].held, l});
    }
    trim.clear();
    if (others + mine <= pol.budget) return true;
    // release idle lanes, largest first, until the batch fits
    std::sort(idle.begin(), idle.end(), [](auto &a, auto &b) { return a.first > b.first; });
    for (auto &p : idle) {
        others -= p.first;
        trim.push_back(p.second);
        if (others + mine <= pol.budget) return true;
    }
    trim.clear();
    return false;
}

inline bool plan_next(const std::deque<Pending> &q, const std::vector<LaneView> &lanes,
                      size_t models_bytes, const Policy &pol, const NeedFn &need_fn, Plan &out) {
    if (q.empty()) return false;
    for (size_t qi = 0; qi < q.size(); qi++) {
        if (qi > 0 && q[0].bypassed >= pol.max_bypass) break;
        const Pending &head = q[qi];
        std::vector<size_t> members{qi};
        int shots = head.shots;
        if (pol.batching)
            for (size_t j = qi + 1; j < q.size(); j++)
                if (q[j].cls == head.cls && shots + q[j].shots <= pol.max_slots) {
                    members.push_back(j);
                    shots += q[j].shots;
                }
        const int lane = pick_lane(lanes, head.cls);
        if (lane < 0) return false;
        while (!members.empty()) {
            int nt = 0;
            shots = 0;
            for (size_t m : members) {
                nt = std::max(nt, q[m].nt);
                shots += q[m].shots;
            }
            size_t need = 0;
            std::vector<int> trim;
            if (fit(lanes, lane, head.cls, shots, nt, models_bytes, pol, need_fn, need, trim)) {
                out.lane = lane;
                out.members = members;
                out.shots = shots;
                out.nt = nt;
                out.need = need;
                out.trim = trim;
                return true;
            }
            members.pop_back();
        }
    }
    return false;
}

}  // namespace serve
}  // namespace inv
