// hz_audit_selftest: checks, on the host alone, that the AsyncProgram audit
// (hz::audit_snapshot, run after launches under HZ_RUN_AUDIT) accepts a
// consistent snapshot of the queues and arena and rejects each of a set of
// planted inconsistencies with the message that names it.
#include <functional>
#include <vector>

#include "common.h"

using hz::capi::guarded;

namespace {

// Two work queues and the free ring, 8 slots each, six claimed links:
// queue 0 holds links 0 (64 promises) and 1 (10), queue 1 holds link 2 (5),
// the free ring holds links 3 and 4, and link 5 is held by a barrier with 7
// promises written. 1000 promises evaluated, so 1000 + 79 + 7 were written.
hz::AsyncSnapshot consistent_snapshot() {
    hz::AsyncSnapshot s;
    s.capacity = 8;
    s.arena_links = 16;
    s.ops = 2;
    s.queues.assign(3, hz::LinkQueue{});
    s.slots.assign(3 * 8, hz::LINK_NONE);
    auto place = [&](unsigned int q, unsigned int head, std::vector<unsigned int> ids) {
        s.queues[q].head = head;
        s.queues[q].tail = head + (unsigned int)ids.size();
        s.queues[q].count = (int)ids.size();
        for (size_t k = 0; k < ids.size(); ++k) s.slots[q * 8 + ((head + k) & 7u)] = ids[k];
    };
    place(0, 5, {0, 1});  // wraps around the ring end
    place(1, 0, {2});
    place(2, 3, {3, 4});
    s.meta = {0u | (64u << 16), 1u | (10u << 16), 0u | (5u << 16), 0, 0, 1u};
    s.aux.assign(12, 0u);
    s.aux[2 * 5 + 1] = (64u << 16) | 7u;
    s.status.claim = 6;
    s.status.units = 3;
    s.status.promises = 1000;
    s.status.appended = 1000 + 79 + 7;
    return s;
}

}  // namespace

extern "C" int hz_audit_selftest(char* report, size_t report_bytes) {
    int missed = -1;
    std::string log;
    int rc = guarded([&] {
        missed = 0;
        hz::AsyncCensus c = hz::audit_snapshot(consistent_snapshot(), true);
        if (c.queued_links != 3 || c.queued_promises != 79 || c.free_links != 2 || c.held_links != 1 ||
            c.held_promises != 7) {
            missed++;
            log += "consistent snapshot: wrong census; ";
        }
        struct Fault {
            const char* name;
            const char* expect;
            bool allow_held;
            std::function<void(hz::AsyncSnapshot&)> plant;
        };
        const std::vector<Fault> faults = {
            {"duplicate id", "appears twice", true, [](hz::AsyncSnapshot& s) { s.slots[1 * 8 + 0] = 1; }},
            {"count off", "tail - head", true, [](hz::AsyncSnapshot& s) { s.queues[1].count = 2; }},
            {"stray slot", "outside [head, tail)", true, [](hz::AsyncSnapshot& s) { s.slots[1 * 8 + 4] = 5; }},
            {"units off", "units", true, [](hz::AsyncSnapshot& s) { s.status.units = 2; }},
            {"lost promise", "unaccounted", true, [](hz::AsyncSnapshot& s) { s.status.appended += 1; }},
            {"leaked link", "neither queued nor free", false, [](hz::AsyncSnapshot&) {}},
            {"bad operation", "names operation", true, [](hz::AsyncSnapshot& s) { s.meta[2] = 7u | (5u << 16); }},
            {"empty queued link", "holds 0 promises", true,
             [](hz::AsyncSnapshot& s) {
                 s.meta[2] = 0u;
                 s.status.appended -= 5;
             }},
            {"group still in", "checked in", true, [](hz::AsyncSnapshot& s) { s.status.checkout = 1; }},
            {"invalid id", "no valid link", true, [](hz::AsyncSnapshot& s) { s.slots[1 * 8 + 0] = 9; }},
        };
        for (const Fault& f : faults) {
            hz::AsyncSnapshot s = consistent_snapshot();
            f.plant(s);
            try {
                hz::audit_snapshot(s, f.allow_held);
                missed++;
                log += std::string(f.name) + ": not caught; ";
            } catch (const std::runtime_error& e) {
                if (std::string(e.what()).find(f.expect) == std::string::npos) {
                    missed++;
                    log += std::string(f.name) + ": caught as '" + e.what() + "'; ";
                }
            }
        }
    });
    if (report && report_bytes) {
        std::snprintf(report, report_bytes, "%s", log.empty() ? "all planted faults caught" : log.c_str());
    }
    return rc != HZ_OK ? rc : missed;
}
