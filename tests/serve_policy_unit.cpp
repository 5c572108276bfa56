// Host-only checks of the request server's admission and batching policy
// (inv.cu_amd/csrc/serve_policy.h).  Built and run by
// tests/test_serve_policy.py; prints one line per check and "ALL OK".
#include <iostream>

#include "../inv.cu_amd/csrc/serve_policy.h"

using namespace inv::serve;

static int failures = 0;
#define CHECK(cond)                                                             \
    do {                                                                        \
        if (cond) std::cout << "ok   " << #cond << "\n";                        \
        else {                                                                  \
            std::cout << "FAIL " << #cond << "  (line " << __LINE__ << ")\n";   \
            failures++;                                                         \
        }                                                                       \
    } while (0)

static Pending req(long long id, const char *cls, int shots, int nt) {
    Pending p;
    p.id = id;
    p.cls = cls;
    p.shots = shots;
    p.nt = nt;
    return p;
}

// need = 100 bytes per shot-step / 1000 of a class's unit size
static size_t need(const std::string &cls, int shots, int nt) {
    const size_t unit = cls == "big" ? 10 : 1;
    return unit * (size_t)shots * (size_t)nt;
}

int main() {
    const NeedFn nf = need;
    Policy pol;
    pol.budget = 100000;
    pol.max_slots = 8;
    pol.max_bypass = 2;

    {  // like-class requests join the oldest in arrival order up to the slot cap
        std::deque<Pending> q{req(1, "a", 3, 1000), req(2, "b", 2, 1000), req(3, "a", 4, 2000),
                              req(4, "a", 2, 500), req(5, "a", 1, 100)};
        std::vector<LaneView> lanes(2);
        Plan p;
        CHECK(plan_next(q, lanes, 0, pol, nf, p));
        CHECK(p.members == (std::vector<size_t>{0, 2, 4}));  // 3 + 4 + 1 = 8 slots; id 4 does not fit
        CHECK(p.shots == 8 && p.nt == 2000);
        CHECK(p.need == need("a", 8, 2000));
        CHECK(p.trim.empty());
    }
    {  // no batching: one request per batch
        Policy one = pol;
        one.batching = false;
        std::deque<Pending> q{req(1, "a", 3, 1000), req(2, "a", 2, 1000)};
        std::vector<LaneView> lanes(1);
        Plan p;
        CHECK(plan_next(q, lanes, 0, one, nf, p) && p.members.size() == 1 && p.shots == 3);
    }
    {  // no idle lane: nothing to do
        std::deque<Pending> q{req(1, "a", 1, 10)};
        std::vector<LaneView> lanes(1);
        lanes[0].state = LANE_BUSY;
        Plan p;
        CHECK(!plan_next(q, lanes, 0, pol, nf, p));
    }
    {  // an idle lane already holding the class is preferred
        std::vector<LaneView> lanes(3);
        lanes[0].cls = "b";
        lanes[0].held = 500;
        lanes[1].cls = "a";
        lanes[1].held = 400;
        CHECK(pick_lane(lanes, "a") == 1);
        CHECK(pick_lane(lanes, "c") == 2);  // the empty lane, not the one holding "b"
    }
    {  // the batch shrinks from its tail to fit the budget
        Policy tight = pol;
        tight.budget = 60000;
        std::deque<Pending> q{req(1, "big", 2, 1000), req(2, "big", 2, 1000), req(3, "big", 2, 1000),
                              req(4, "big", 2, 1000)};
        std::vector<LaneView> lanes(2);
        Plan p;
        CHECK(plan_next(q, lanes, 5000, tight, nf, p));
        CHECK(p.members == (std::vector<size_t>{0, 1}));  // 40000 + 5000 <= 60000; 60000 + 5000 is not
    }
    {  // idle lanes are released (largest first) to make room
        std::deque<Pending> q{req(1, "big", 8, 1000)};  // needs 80000
        std::vector<LaneView> lanes(3);
        lanes[0].cls = "x";
        lanes[0].held = 30000;
        lanes[1].cls = "y";
        lanes[1].held = 20000;
        lanes[2].cls = "z";
        lanes[2].held = 5000;
        Plan p;
        CHECK(plan_next(q, lanes, 0, pol, nf, p));
        CHECK(p.lane == 2);                       // the idle lane holding least
        CHECK(p.trim == (std::vector<int>{0}));  // 5000 + 80000 fits once lane 0 goes
    }
    {  // busy lanes are never released; the head waits and younger ones backfill
        std::deque<Pending> q{req(1, "big", 8, 1000), req(2, "a", 2, 1000)};
        std::vector<LaneView> lanes(2);
        lanes[0].state = LANE_BUSY;
        lanes[0].cls = "a";
        lanes[0].held = 50000;
        Plan p;
        CHECK(plan_next(q, lanes, 0, pol, nf, p));
        CHECK(p.members == (std::vector<size_t>{1}) && p.lane == 1);
        q[0].bypassed = 2;  // overtaken max_bypass times: no more backfill
        CHECK(!plan_next(q, lanes, 0, pol, nf, p));
    }
    {  // a lane keeping more than a batch needs reserves what it keeps
        std::deque<Pending> q{req(1, "a", 1, 10)};
        std::vector<LaneView> lanes(1);
        lanes[0].cls = "a";
        lanes[0].held = 99990;
        Plan p;
        CHECK(plan_next(q, lanes, 0, pol, nf, p));
        CHECK(plan_next(q, lanes, 20, pol, nf, p) == false);  // 99990 + 20 > budget, nothing to trim
    }
    std::cout << (failures ? "FAILED" : "ALL OK") << "\n";
    return failures ? 1 : 0;
}
