// hbt_server.hip -- request server for independent HBT analyses on one GPU
// (C ABI in include/hbt_hip.h, "request server").
//
// A request is split into parts of whole groups (about part_particles input
// particles each).  Each part is prepared on the host, uploaded and launched
// on its own (its own arena block and view), adding to the request's one
// histogram block, so a large request's host preparation, upload and
// launches overlap part by part instead of running one after the other.
//
// Threads (each owns one stream, all at the default priority):
//   caller     hbt_server_submit: counts the kept particles and the pairs
//              (pair-cap and footprint checks), draws the mixing plan, then
//              per part: rapidity cut, prefilter records, work items on the
//              host, and queues the part.
//   admission  takes queued parts in arrival order among those that fit
//              (device arena block + view slot; with a request's first part
//              also its histogram block and result slot), uploads them on
//              the upload stream, marks them active.
//   launcher   each launch takes up to step_pairs enumerated pairs of work
//              items from the active parts (water-filling, smallest
//              remaining first), copies the items and launches k_fast on
//              the compute stream; at most two launches are in flight.  A
//              part whose last items went into a launch is retired after
//              that launch; the launch that carries a request's last items
//              is followed by the copy of its histograms to its pinned
//              result slot on the copy stream.
//   completion waits for those launches and copies in order, frees the
//              parts' arena blocks and views, and with a request's last part
//              its histogram block; stamps t_done.
//
// HBM layout: a request's histogram block (bins [nbins][8] doubles, the four
// q_inv arrays (count, den, sum, num) of nq words each, the deferred-pair
// count -- copied back in one piece -- and the deferred-pair records), and
// per part (256-byte aligned parts): particle records (64 B), prefilter
// records (32 or 16 B), c/E bounds, plan windows, (cos, sin) per row, and
// psi_ref per group.  A part's same-event list also serves as its mixed list
// (partners are events of the same group).
#include <hip/hip_runtime.h>

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "hbt_internal.h"

using namespace hbt;

namespace {

using Clock = std::chrono::steady_clock;

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

constexpr size_t kHostPool = 4;  // host lists kept for reuse

// First-fit allocator over the device arena.
class Arena {
  public:
    void init(size_t size) {
        free_.clear();
        if (size) free_[0] = size;
        total_ = size;
        used_ = peak_ = 0;
    }
    bool alloc(size_t n, size_t &off) {
        n = align256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            off = it->first;
            const size_t rest = it->second - n, start = it->first + n;
            free_.erase(it);
            if (rest) free_[start] = rest;
            used_ += n;
            peak_ = std::max(peak_, used_);
            return true;
        }
        return false;
    }
    void release(size_t off, size_t n) {
        n = align256(n);
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    size_t total() const { return total_; }
    size_t used() const { return used_; }
    size_t peak() const { return peak_; }

  private:
    std::map<size_t, size_t> free_;  // offset -> size
    size_t total_ = 0, used_ = 0, peak_ = 0;
};

struct Request;

// Consecutive whole groups of a request, resident and launched on their own.
struct Part {
    Request *req = nullptr;
    int e0 = 0, nev = 0, ngroups = 0;  // first event and counts (request indices)
    HostParticles hp;
    std::vector<int> grp_ev;           // part-local group offsets
    std::vector<long long> rowoff;     // plan rows per part-local event
    std::vector<int> plan_ids;         // partner: part-local event index
    std::vector<double> plan_cs;
    std::vector<WorkItem> items;
    std::vector<JWin> wins;  // partner windows of the mixed items
    std::vector<uint64_t> item_pairs;
    size_t next = 0;           // next item to launch
    uint64_t pairs_left = 0;   // enumerated pairs not yet launched
    double pmax2 = 0.0;
    // device residency
    size_t foot = 0, arena_off = 0;
    size_t o_rec = 0, o_srec = 0, o_pf = 0, o_cw = 0, o_spf = 0, o_scw = 0, o_orig = 0, o_win = 0,
           o_cs = 0, o_psi = 0;
    int view = -1;
    View v{};
    bool zero_hist = false;  // the request's first part admitted: clears its histograms
    bool last = false;       // carries the request's last items: its result is copied after
    bool ev_ok = false;      // done_ev recorded
    hipEvent_t done_ev = nullptr;
};

struct Request {
    int64_t id = 0;
    int state = HBT_REQ_QUEUED;
    std::string err;
    int nev = 0, ngroups = 0;
    int nparts = 0;          // parts of the split (fixed at submit)
    int parts_admitted = 0, parts_finished = 0;
    int64_t n_kept = 0;
    uint64_t enum_same = 0, enum_mixed = 0;
    size_t foot = 0;                         // histogram block + the parts' blocks
    size_t hist_foot = 0, hist_off = 0;      // histogram block in the arena
    bool hist_alloc = false;
    int rslot = -1;
    std::vector<std::unique_ptr<Part>> parts;
    double t_submit = 0, t_queued = 0, t_admit = 0, t_active = 0, t_first = 0, t_last = 0,
           t_done = 0;
};

}  // namespace

struct hbt_server {
    hbt_config cfg;
    Derived D;
    hbt_server_config sc;
    ItemShape shape;
    bool f32 = false, ext = false;
    int device = 0, cu_count = 0;
    hipStream_t s_compute = nullptr, s_copy = nullptr, s_upload = nullptr;
    char *d_arena = nullptr;
    Arena arena;
    View *d_views = nullptr;
    unsigned long long *d_misc = nullptr;  // [0..9] stage words, [16..17] binned counters
    int *d_counter = nullptr;
    std::vector<int> free_views;
    std::vector<char *> h_results;
    std::vector<int> free_results;
    size_t result_bytes = 0;
    size_t defer_count_off = 0;  // byte offset of the deferred-pair count in a result
    long long part_particles = 0;  // input particles per part of a request
    // host lists of uploaded parts, kept for the next parts' gathers (their
    // pages are faulted in already); at most kHostPool
    std::vector<HostParticles> host_pool;
    WorkItem *d_items[2] = {};
    WorkItem *h_items[2] = {};
    char *h_stage = nullptr;  // pinned staging of uploads (admission thread only)
    size_t stage_cap = 0;
    hipEvent_t step_ev[2] = {};     // end of launch (ordering)
    hipEvent_t step_t0[2] = {};     // start of launch (timing)
    hipEvent_t step_t1[2] = {};     // end of launch (timing)
    bool step_pending[2] = {};
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Part *> queue;      // prepared parts, arrival order
    std::vector<Part *> active;    // resident parts with items left
    std::deque<Part *> draining;   // parts whose last launch (and copy) is queued
    std::map<int64_t, std::unique_ptr<Request>> reqs;
    std::thread t_admission, t_launcher, t_completion;
    bool stop = false;
    bool launcher_done = false;  // no more readbacks will be queued
    int64_t next_id = 1;
    hbt_server_stats st{};
    Clock::time_point t0;
    std::string err;

    double now() const {
        return std::chrono::duration<double>(Clock::now() - t0).count();
    }
};

namespace {

std::string g_server_create_err;

int sfail(hbt_server *s, int code, const std::string &msg) {
    if (s) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = msg;
    } else {
        g_server_create_err = msg;
    }
    return code;
}

// A request's histogram block: the result (copied back in one piece), then
// its deferred-pair records.
size_t hist_layout(const hbt_server *s) {
    return align256(s->result_bytes) + align256((size_t)kReqDeferCap * sizeof(DeferRec));
}

// Part layout in its arena block for nf fast-list particles, rows plan rows,
// nwins partner windows and ngroups groups; returns the footprint.
size_t part_layout(const hbt_server *s, Part *p, long long nf, long long rows, size_t nwins,
                   int ngroups) {
    Part tmp;
    Part &r = p ? *p : tmp;
    size_t o = 0;
    // fast-list records and their cell-sorted copy, the prefilter records of both
    for (size_t *po : {&r.o_rec, &r.o_srec}) {
        *po = o;
        o = align256(o + (size_t)nf * sizeof(PartRec));
    }
    for (size_t *po : {&r.o_pf, &r.o_spf}) {
        *po = o;
        o = align256(o + (size_t)nf * (s->f32 ? sizeof(PfRec32) : sizeof(PfRec64)));
    }
    for (size_t *po : {&r.o_cw, &r.o_scw}) {
        *po = o;
        o = align256(o + (size_t)nf * (s->f32 ? 4 : 8));
    }
    r.o_orig = o;
    o = align256(o + (size_t)nf * sizeof(OrigRank));
    r.o_win = o;
    o = align256(o + std::max<size_t>(nwins, 1) * sizeof(JWin));
    r.o_cs = o;
    o = align256(o + (size_t)std::max<long long>(rows, 1) * 16);
    r.o_psi = o;
    o = align256(o + (size_t)std::max(ngroups, 1) * 8);
    return o;
}

void make_part_view(hbt_server *s, Part &p) {
    const Request &r = *p.req;
    char *h = s->d_arena + r.hist_off;
    char *b = s->d_arena + p.arena_off;
    View &v = p.v;
    v.H.bins = (double *)h;
    unsigned long long *q = (unsigned long long *)(h + (size_t)s->D.nbins * 64);
    v.H.qinv_count = q;
    v.H.qinv_den = q + s->D.nq;
    v.H.qinv_sum = (double *)(q + 2 * s->D.nq);
    v.H.qinv_num = (double *)(q + 3 * s->D.nq);
    v.H.stage = s->d_misc;
    // pairs left to the host: records in the histogram block, their count in
    // the zeroed result tail, copied back with the histograms
    v.H.defer = (DeferRec *)(h + align256(s->result_bytes));
    v.H.ndefer = (unsigned *)(h + s->defer_count_off);
    v.H.defer_cap = kReqDeferCap;
    v.A = v.B = (const PartRec *)(b + p.o_rec);
    v.pfA = v.pfB = (const void *)(b + p.o_pf);
    v.cwA = v.cwB = (const void *)(b + p.o_cw);
    v.pfS = (const void *)(b + p.o_spf);
    v.cwS = (const void *)(b + p.o_scw);
    v.origS = (const OrigRank *)(b + p.o_orig);
    v.AS = (const PartRec *)(b + p.o_srec);
    v.mwin = (const JWin *)(b + p.o_win);
    v.plan_cs = (const double *)(b + p.o_cs);
    v.psi = (const double *)(b + p.o_psi);
    v.P = s->D.P;
}

// Uploads a part into its arena block on the upload stream (clearing the
// request's histograms first when it is the request's first part) and
// waits.
hipError_t upload_part(hbt_server *s, Part &p) {
    char *b = s->d_arena + p.arena_off;
    const long long nf = p.hp.nf;
    hipError_t e = hipSuccess;
    if (p.zero_hist)
        e = launch_zero_words(s->d_arena + p.req->hist_off, s->result_bytes / 8, s->s_upload);
    std::vector<StagePiece> pieces;
    auto cp = [&](size_t off, const void *src, size_t bytes) {
        if (bytes) pieces.push_back(StagePiece{b + off, src, bytes});
    };
    const std::pair<const RecordSet *, size_t> sets[] = {{&p.hp.fast, 0}, {&p.hp.srt, 1}};
    for (const auto &x : sets) {
        const RecordSet &R = *x.first;
        const bool srt = x.second == 1;
        if (!srt) cp(p.o_rec, R.rec.data(), (size_t)nf * sizeof(PartRec));
        if (s->f32) {
            cp(srt ? p.o_spf : p.o_pf, R.pf32.data(), (size_t)nf * sizeof(PfRec32));
            cp(srt ? p.o_scw : p.o_cw, R.cw32.data(), (size_t)nf * 4);
        } else {
            cp(srt ? p.o_spf : p.o_pf, R.pf64.data(), (size_t)nf * sizeof(PfRec64));
            cp(srt ? p.o_scw : p.o_cw, R.cw64.data(), (size_t)nf * 8);
        }
    }
    cp(p.o_orig, p.hp.orr.data(), p.hp.orr.size() * sizeof(OrigRank));
    cp(p.o_win, p.wins.data(), p.wins.size() * sizeof(JWin));
    cp(p.o_cs, p.plan_cs.data(), p.plan_cs.size() * 8);
    cp(p.o_psi, p.hp.psi2.data(), p.hp.psi2.size() * 8);
    // the previous part's copies out of the staging buffer have completed
    // (its upload ended with a synchronization of this stream)
    if (e == hipSuccess) e = ensure_pinned(&s->h_stage, &s->stage_cap, stage_bytes(pieces));
    if (e == hipSuccess) e = stage_pieces(s->h_stage, pieces, s->s_upload);
    if (e == hipSuccess)
        e = gather_srt((const PartRec *)(b + p.o_rec), (const OrigRank *)(b + p.o_orig),
                       (PartRec *)(b + p.o_srec), nf, s->s_upload);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s->d_views + p.view, &p.v, sizeof(View), hipMemcpyHostToDevice,
                           s->s_upload);
    if (e == hipSuccess) e = hipStreamSynchronize(s->s_upload);
    return e;
}

// Frees what a part holds on the device.  Caller holds the lock.
void retire_part(hbt_server *s, Part &p) {
    if (p.foot) s->arena.release(p.arena_off, p.foot);
    p.foot = 0;
    if (p.view >= 0) {
        s->free_views.push_back(p.view);
        s->st.resident--;
    }
    p.view = -1;
}

// Frees a finished request's histogram block.  Caller holds the lock.
void retire_request(hbt_server *s, Request &r) {
    if (r.hist_alloc) s->arena.release(r.hist_off, r.hist_foot);
    r.hist_alloc = false;
}

void admission_loop(hbt_server *s) {
    hipSetDevice(s->device);
    for (;;) {
        Part *p = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            for (;;) {
                if (s->stop) return;
                // first part, in arrival order, that fits: its block, and with
                // its request's first part the histogram block and a result slot
                if (!s->free_views.empty()) {
                    for (auto it = s->queue.begin(); it != s->queue.end(); ++it) {
                        Request &r = *(*it)->req;
                        const bool first = !r.hist_alloc;
                        if (first && s->free_results.empty()) continue;
                        size_t hoff = 0, off = 0;
                        if (first && !s->arena.alloc(r.hist_foot, hoff)) continue;
                        if (!s->arena.alloc((*it)->foot, off)) {
                            if (first) s->arena.release(hoff, r.hist_foot);
                            continue;
                        }
                        p = *it;
                        p->arena_off = off;
                        if (first) {
                            r.hist_off = hoff;
                            r.hist_alloc = true;
                            r.rslot = s->free_results.back();
                            s->free_results.pop_back();
                            r.state = HBT_REQ_LOADING;
                            r.t_admit = s->now();
                            p->zero_hist = true;
                        }
                        s->queue.erase(it);
                        break;
                    }
                }
                if (p) break;
                s->cv.wait(lk);
            }
            p->view = s->free_views.back();
            s->free_views.pop_back();
            p->req->parts_admitted++;
            s->st.resident++;
            s->st.resident_peak = std::max(s->st.resident_peak, s->st.resident);
        }
        make_part_view(s, *p);
        const hipError_t e = upload_part(s, *p);
        std::lock_guard<std::mutex> lk(s->mu);
        Request &r = *p->req;
        if (e != hipSuccess) {
            // no launches for this part; it still finishes and retires in
            // order, and the request ends failed
            if (r.err.empty()) r.err = std::string("upload: ") + hipGetErrorString(e);
            p->next = p->items.size();
            p->pairs_left = 0;
        }
        if (r.state == HBT_REQ_LOADING) {
            r.state = HBT_REQ_ACTIVE;
            r.t_active = s->now();
        }
        // host copies are on the device now: the lists go back to the pool
        HostParticles keep_meta;
        keep_meta.n = p->hp.n;
        keep_meta.n_in = p->hp.n_in;
        if (s->host_pool.size() < kHostPool) s->host_pool.push_back(std::move(p->hp));
        p->hp = std::move(keep_meta);
        std::vector<double>().swap(p->plan_cs);
        std::vector<JWin>().swap(p->wins);
        s->active.push_back(p);
        s->cv.notify_all();
    }
}

// One launch: items of the active parts, water-filling over parts ordered
// by remaining pairs.  Caller holds the lock.  Parts whose last items are
// taken move to `finishing`; the one that completes its request is marked
// last.
void build_step(hbt_server *s, std::vector<WorkItem> &items, std::vector<Part *> &finishing,
                std::vector<Part *> &launched, double &pmax2, double &pairs) {
    std::vector<Part *> act = s->active;
    std::sort(act.begin(), act.end(), [](const Part *a, const Part *b) {
        return a->pairs_left < b->pairs_left;
    });
    double budget = s->sc.step_pairs;
    const size_t cap = (size_t)s->sc.max_step_items;
    pmax2 = 0.0;
    pairs = 0.0;
    for (size_t k = 0; k < act.size() && items.size() < cap; k++) {
        Part *r = act[k];
        const double share = budget / (double)(act.size() - k);
        double took = 0.0;
        bool any = false;
        while (r->next < r->items.size() && items.size() < cap && (!any || took < share)) {
            WorkItem w = r->items[r->next];
            w.view = r->view;
            items.push_back(w);
            took += (double)r->item_pairs[r->next];
            r->pairs_left -= r->item_pairs[r->next];
            r->next++;
            any = true;
        }
        budget = std::max(0.0, budget - took);
        pairs += took;
        if (any) {
            pmax2 = std::max(pmax2, r->pmax2);
            launched.push_back(r);
        }
        if (r->next == r->items.size()) finishing.push_back(r);
    }
    for (Part *p : finishing) {
        Request &r = *p->req;
        if (++r.parts_finished == r.nparts) {
            p->last = true;
            r.state = HBT_REQ_DRAINING;
        }
        s->active.erase(std::find(s->active.begin(), s->active.end(), p));
    }
}

// Waits for launch buffer b (if in flight) and adds its device time.
hipError_t retire_launch(hbt_server *s, int b) {
    if (!s->step_pending[b]) return hipSuccess;
    hipError_t e = hipEventSynchronize(s->step_ev[b]);
    s->step_pending[b] = false;
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, s->step_t0[b], s->step_t1[b]);
    std::lock_guard<std::mutex> lk(s->mu);
    s->st.busy_s += 1e-3 * ms;
    return e;
}

void launcher_loop(hbt_server *s) {
    hipSetDevice(s->device);
    std::vector<WorkItem> items;
    std::vector<Part *> finishing, launched;
    int step = 0;
    for (;;) {
        items.clear();
        finishing.clear();
        launched.clear();
        double pmax2 = 0.0, pairs = 0.0;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->active.empty() && !s->stop && (s->step_pending[0] || s->step_pending[1])) {
                // idle: account the launches still in flight first
                lk.unlock();
                retire_launch(s, 0);
                retire_launch(s, 1);
                lk.lock();
            }
            s->cv.wait(lk, [&] { return s->stop || !s->active.empty(); });
            if (s->stop) break;
            build_step(s, items, finishing, launched, pmax2, pairs);
        }
        const int buf = step & 1;
        hipError_t e = retire_launch(s, buf);
        if (!items.empty() && e == hipSuccess) {
            std::memcpy(s->h_items[buf], items.data(), items.size() * sizeof(WorkItem));
            e = hipMemcpyAsync(s->d_items[buf], s->h_items[buf], items.size() * sizeof(WorkItem),
                               hipMemcpyHostToDevice, s->s_compute);
            if (e == hipSuccess) e = hipEventRecord(s->step_t0[buf], s->s_compute);
            FastArgs a;
            a.slab = SlabArgs{};  // the server adds binned pairs with atomics
            a.P = s->D.P;
            prefilter_margins(s->D, pmax2, a.P);
            a.views = s->d_views;
            a.items = s->d_items[buf];
            a.nitems = (int)items.size();
            a.nqueues = 1;  // one queue: a step mixes the items of several requests
            for (int q = 0; q <= kMaxQueues; q++) a.qoff[q] = q == 0 ? 0 : a.nitems;
            a.counter = s->d_counter;
            a.binned = s->d_misc + 16;
            const long blocks = std::max(
                1l, std::min<long>((long)s->cu_count * s->sc.tune.blocks_per_cu,
                                   ((long)items.size() + 3) / 4));
            if (e == hipSuccess)
                e = launch_fast_kernel(a, s->f32, s->sc.tune.R, s->ext, blocks, s->s_compute);
            if (e == hipSuccess) e = hipEventRecord(s->step_t1[buf], s->s_compute);
            if (e == hipSuccess) e = hipEventRecord(s->step_ev[buf], s->s_compute);
            if (e == hipSuccess) {
                s->step_pending[buf] = true;
                const double tl = s->now();
                std::lock_guard<std::mutex> lk(s->mu);
                s->st.launches++;
                s->st.items += (int64_t)items.size();
                s->st.pairs += pairs;
                for (Part *p : launched) {
                    if (p->req->t_first == 0) p->req->t_first = tl;
                    p->req->t_last = tl;
                }
            }
            step++;
        }
        // parts whose items are all launched retire once this launch has
        // finished; the request's last part also has its histograms copied
        // back after it (a part that never had items needs only its upload,
        // which has completed)
        for (Part *p : finishing) {
            Request &r = *p->req;
            hipError_t e2 = e;
            if (e2 == hipSuccess && !items.empty()) {
                e2 = hipStreamWaitEvent(s->s_copy, s->step_ev[buf], 0);
            } else {
                // no launch this step: the launches still in flight may carry
                // the request's other parts
                for (int b2 = 0; b2 < 2 && e2 == hipSuccess; b2++)
                    if (s->step_pending[b2]) e2 = hipStreamWaitEvent(s->s_copy, s->step_ev[b2], 0);
            }
            if (e2 == hipSuccess && p->last)
                e2 = hipMemcpyAsync(s->h_results[r.rslot], s->d_arena + r.hist_off,
                                    s->result_bytes, hipMemcpyDeviceToHost, s->s_copy);
            if (e2 == hipSuccess) e2 = hipEventRecord(p->done_ev, s->s_copy);
            std::lock_guard<std::mutex> lk(s->mu);
            p->ev_ok = e2 == hipSuccess;
            if (e2 != hipSuccess && r.err.empty())
                r.err = std::string("launch/readback: ") + hipGetErrorString(e2);
            s->draining.push_back(p);
            s->cv.notify_all();
        }
    }
    // drain what is in flight before the streams go away
    for (int b = 0; b < 2; b++) retire_launch(s, b);
    std::lock_guard<std::mutex> lk(s->mu);
    s->launcher_done = true;
    s->cv.notify_all();
}

// Pairs the device left to the host for a finished request (K_phi at a bin
// edge): decided with glibc's atan2 (resolve_records) and added to the
// request's result in host memory, before it is reported done.
hipError_t resolve_request_defer(hbt_server *s, Request &r, std::string &err) {
    char *res = s->h_results[r.rslot];
    unsigned n = 0;
    std::memcpy(&n, res + s->defer_count_off, sizeof(n));
    if (n == 0) return hipSuccess;
    if (n > kReqDeferCap) {
        err = "more K_phi bin-edge pairs than the request's deferral buffer";
        return hipSuccess;
    }
    std::vector<DeferRec> rec(n);
    const hipError_t e = hipMemcpy(rec.data(), s->d_arena + r.hist_off + align256(s->result_bytes),
                                   n * sizeof(DeferRec), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    std::vector<DeferAdd> adds;
    resolve_records(s->D.P, rec.data(), n, adds);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->st.deferred += (int32_t)n;
    }
    double *bins = (double *)res;
    for (const DeferAdd &a : adds) {
        double *b = bins + a.k * BIN_STRIDE;
        if (a.kind == 0) {
            b[BIN_COUNT] += 1.0;
            b[BIN_QO] += a.v[1];
            b[BIN_QS] += a.v[2];
            b[BIN_QL] += a.v[3];
            b[BIN_NUM] += a.v[4];
        } else {
            unsigned long long den;
            std::memcpy(&den, b + BIN_DEN, 8);
            den += 1;
            std::memcpy(b + BIN_DEN, &den, 8);
        }
    }
    return hipSuccess;
}

void completion_loop(hbt_server *s) {
    hipSetDevice(s->device);
    for (;;) {
        Part *p = nullptr;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv.wait(lk, [&] { return s->launcher_done || !s->draining.empty(); });
            if (s->draining.empty()) return;  // stopped and nothing left to read back
            p = s->draining.front();
        }
        Request &r = *p->req;
        hipError_t e = p->ev_ok ? hipEventSynchronize(p->done_ev) : hipSuccess;
        std::string derr;
        bool failed;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            failed = !r.err.empty();
        }
        // the request's other parts have retired before its last one: they
        // finished in earlier launches or earlier in the same one
        if (p->last && e == hipSuccess && !failed) e = resolve_request_defer(s, r, derr);
        std::lock_guard<std::mutex> lk(s->mu);
        s->draining.pop_front();
        retire_part(s, *p);
        if (e != hipSuccess && r.err.empty()) r.err = std::string("readback: ") + hipGetErrorString(e);
        if (!derr.empty() && r.err.empty()) r.err = derr;
        if (p->last) {
            retire_request(s, r);
            r.t_done = s->now();
            if (r.err.empty()) {
                r.state = HBT_REQ_DONE;
                s->st.completed++;
            } else {
                r.state = HBT_REQ_FAILED;
                s->st.failed++;
            }
        }
        s->cv.notify_all();
    }
}

Request *find(hbt_server *s, int64_t id) {
    auto it = s->reqs.find(id);
    return it == s->reqs.end() ? nullptr : it->second.get();
}

void destroy_resources(hbt_server *s) {
    if (s->s_compute) hipStreamSynchronize(s->s_compute);
    if (s->s_copy) hipStreamSynchronize(s->s_copy);
    if (s->s_upload) hipStreamSynchronize(s->s_upload);
    for (auto &kv : s->reqs)
        for (auto &p : kv.second->parts)
            if (p->done_ev) hipEventDestroy(p->done_ev);
    for (char *p : s->h_results)
        if (p) hipHostFree(p);
    if (s->h_stage) hipHostFree(s->h_stage);
    for (int b = 0; b < 2; b++) {
        if (s->h_items[b]) hipHostFree(s->h_items[b]);
        if (s->d_items[b]) hipFree(s->d_items[b]);
        if (s->step_ev[b]) hipEventDestroy(s->step_ev[b]);
        if (s->step_t0[b]) hipEventDestroy(s->step_t0[b]);
        if (s->step_t1[b]) hipEventDestroy(s->step_t1[b]);
    }
    if (s->d_views) hipFree(s->d_views);
    if (s->d_misc) hipFree(s->d_misc);
    if (s->d_counter) hipFree(s->d_counter);
    if (s->d_arena) hipFree(s->d_arena);
    if (s->s_compute) hipStreamDestroy(s->s_compute);
    if (s->s_copy) hipStreamDestroy(s->s_copy);
    if (s->s_upload) hipStreamDestroy(s->s_upload);
}

}  // namespace

extern "C" {

int hbt_server_create(const hbt_config *cfg, const hbt_server_config *sc_in, hbt_server **out) {
    if (!cfg || !out) return sfail(nullptr, HBT_ERR_ARG, "null argument");
    *out = nullptr;
    if (cfg->kernel != HBT_KERNEL_FAST)
        return sfail(nullptr, HBT_ERR_ARG, "the server runs the fast kernel (kernel = 1)");
    std::unique_ptr<hbt_server> s(new hbt_server());
    std::string err;
    int rc = derive(*cfg, s->D, err);
    if (rc) return sfail(nullptr, rc, err);
    s->cfg = *cfg;
    hbt_server_config sc{};
    if (sc_in) {
        sc = *sc_in;
    } else {
        sc.hbm_bytes = 16ll << 30;
        sc.max_resident = 16;
        sc.max_step_items = 1 << 18;
        sc.step_pairs = 2e9;
        sc.tune = hbt_tuning{2, 1024, 4, 8, 0.0};
        sc.part_particles = 0;
    }
    const hbt_tuning &t = sc.tune;
    if ((t.R != 1 && t.R != 2 && t.R != 4) || t.jchunk < 64 || t.jchunk > (1 << 24) ||
        t.kchunk < 1 || t.kchunk > 64 || t.blocks_per_cu < 1 || t.blocks_per_cu > 64)
        return sfail(nullptr, HBT_ERR_ARG, "tuning out of range");
    if (sc.hbm_bytes < (1 << 20) || sc.max_resident < 1 || sc.max_resident > 4096 ||
        sc.max_step_items < 1 || !(sc.step_pairs > 0) || sc.part_particles < 0)
        return sfail(nullptr, HBT_ERR_ARG, "bad server configuration");
    s->sc = sc;
    s->part_particles = sc.part_particles > 0 ? sc.part_particles : 1000000;
    s->shape = item_shape(sc.tune);
    s->f32 = cfg->precision == HBT_PREC_FP32C;
    s->ext = s->D.P.qinv || s->D.P.azimuthal;
    s->defer_count_off = (size_t)s->D.nbins * 64 + (size_t)s->D.nq * 32;
    s->result_bytes = s->defer_count_off + 256;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return sfail(nullptr, HBT_ERR_NODEVICE, "no HIP device visible");
    if (cfg->device < 0 || cfg->device >= ndev)
        return sfail(nullptr, HBT_ERR_NODEVICE, "bad device");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, cfg->device) != hipSuccess)
        return sfail(nullptr, HBT_ERR_NODEVICE, "hipGetDeviceProperties failed");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return sfail(nullptr, HBT_ERR_NODEVICE,
                     std::string("built for gfx950, device is ") + prop.gcnArchName);
    s->device = cfg->device;
    s->cu_count = prop.multiProcessorCount;
    if (hipSetDevice(s->device) != hipSuccess)
        return sfail(nullptr, HBT_ERR_NODEVICE, "hipSetDevice failed");
    auto bad = [&](hipError_t e, const char *what) {
        if (e == hipSuccess) return false;
        g_server_create_err = std::string(what) + ": " + hipGetErrorString(e);
        return true;
    };
    bool ok = !bad(hipStreamCreateWithFlags(&s->s_compute, hipStreamNonBlocking), "stream") &&
              !bad(hipStreamCreateWithFlags(&s->s_copy, hipStreamNonBlocking), "stream") &&
              !bad(hipStreamCreateWithFlags(&s->s_upload, hipStreamNonBlocking), "stream");
    if (ok && bad(hipMalloc((void **)&s->d_arena, (size_t)sc.hbm_bytes), "arena hipMalloc")) {
        s->d_arena = nullptr;
        ok = false;
    }
    ok = ok && !bad(hipMalloc((void **)&s->d_views, sizeof(View) * sc.max_resident), "hipMalloc");
    ok = ok && !bad(hipMalloc((void **)&s->d_misc, 32 * 8), "hipMalloc");
    ok = ok && !bad(hipMalloc((void **)&s->d_counter, 4), "hipMalloc");
    ok = ok && !bad(hipMemset(s->d_misc, 0, 32 * 8), "hipMemset");
    for (int b = 0; b < 2 && ok; b++) {
        ok = !bad(hipMalloc((void **)&s->d_items[b], sizeof(WorkItem) * sc.max_step_items),
                  "hipMalloc") &&
             !bad(hipHostMalloc((void **)&s->h_items[b], sizeof(WorkItem) * sc.max_step_items),
                  "hipHostMalloc") &&
             !bad(hipEventCreateWithFlags(&s->step_ev[b], hipEventDisableTiming), "event") &&
             !bad(hipEventCreate(&s->step_t0[b]), "event") &&
             !bad(hipEventCreate(&s->step_t1[b]), "event");
    }
    s->h_results.assign(sc.max_resident, nullptr);
    for (int k = 0; k < sc.max_resident && ok; k++)
        ok = !bad(hipHostMalloc((void **)&s->h_results[k], s->result_bytes), "hipHostMalloc");
    if (!ok) {
        const int code = s->d_arena ? HBT_ERR_HIP : HBT_ERR_NOMEM;
        destroy_resources(s.get());
        return code;
    }
    s->arena.init((size_t)sc.hbm_bytes);
    for (int k = sc.max_resident - 1; k >= 0; k--) {
        s->free_views.push_back(k);
        s->free_results.push_back(k);
    }
    s->st.arena_bytes = sc.hbm_bytes;
    s->t0 = Clock::now();
    hbt_server *sp = s.get();
    sp->t_admission = std::thread(admission_loop, sp);
    sp->t_launcher = std::thread(launcher_loop, sp);
    sp->t_completion = std::thread(completion_loop, sp);
    *out = s.release();
    return HBT_OK;
}

int hbt_server_destroy(hbt_server *s) {
    if (!s) return HBT_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    // the admission thread may be inside an upload and the launcher inside a
    // launch: each ends its current piece of work, then sees stop
    if (s->t_admission.joinable()) s->t_admission.join();
    if (s->t_launcher.joinable()) s->t_launcher.join();
    if (s->t_completion.joinable()) s->t_completion.join();
    destroy_resources(s);
    delete s;
    return HBT_OK;
}

const char *hbt_server_last_error(const hbt_server *s) {
    return s ? s->err.c_str() : g_server_create_err.c_str();
}

int hbt_server_submit(hbt_server *s, int32_t nev, const int64_t *ev_off,
                      const double *const soa[8], int32_t events_per_group, int32_t seed,
                      int64_t *req_id) {
    if (!s || !req_id) return HBT_ERR_ARG;
    if (nev < 1 || events_per_group < 1)
        return sfail(s, HBT_ERR_ARG, "need nev >= 1 and events_per_group >= 1");
    if (!ev_off || !soa) return sfail(s, HBT_ERR_ARG, "null particle arrays");
    for (int a = 0; a < 8; a++)
        if (!soa[a]) return sfail(s, HBT_ERR_ARG, "null particle column");
    for (int e = 0; e < nev; e++)
        if (ev_off[e + 1] < ev_off[e]) return sfail(s, HBT_ERR_ARG, "event offsets decrease");
    std::unique_ptr<Request> r(new Request());
    r->t_submit = s->now();
    r->nev = nev;
    r->ngroups = (nev + events_per_group - 1) / events_per_group;
    std::vector<int> grp_ev(r->ngroups + 1);
    for (int g = 0; g <= r->ngroups; g++) grp_ev[g] = std::min(nev, g * events_per_group);
    // kept particles per event (the gathers' rapidity cut): the pair totals
    // for the pair-cap check and the footprint, before any part is queued
    const double cut_max = tanh(s->cfg.HBTrap_max), cut_min = tanh(s->cfg.HBTrap_min);
    std::vector<long long> off((size_t)nev + 1, 0);
    const long long n_in = ev_off[nev] - ev_off[0];
    parallel_ranges(nev, std::max(1ll, 2048ll * nev / std::max(1ll, n_in)),
                    [&](long long e0, long long e1) {
        for (long long e = e0; e < e1; e++) {
            long long c = 0;
            for (int64_t i = ev_off[e]; i < ev_off[e + 1]; i++)
                c += rapidity_kept(soa[2][i], soa[3][i], cut_min, cut_max) ? 1 : 0;
            off[(size_t)e + 1] = c;
        }
    });
    for (int e = 0; e < nev; e++) {
        if (off[(size_t)e + 1] >= (1 << 18))
            return sfail(s, HBT_ERR_CAPACITY, "an event keeps more than 262143 particles");
        off[(size_t)e + 1] += off[(size_t)e];
    }
    r->n_kept = off[(size_t)nev];
    // mixing plan, buffer after buffer from one stream (HBT_correlation.cpp:200-216)
    std::mt19937 gen((unsigned)seed);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    std::uniform_int_distribution<int> uint_{};
    std::vector<long long> rowoff((size_t)nev + 1, 0);
    std::vector<int> plan_ids;  // partner: event index of the request
    std::vector<double> plan_cs;
    for (int g = 0; g < r->ngroups; g++) {
        const int e0 = grp_ev[g], ne = grp_ev[g + 1] - e0;
        const int M = s->cfg.number_of_mixed_events > 0 ? s->cfg.number_of_mixed_events
                                                        : ne / 2 + 1;
        for (int iev = 0; iev < ne; iev++) {
            const size_t base = plan_ids.size();
            for (int count = 0; count < M; count++) {
                int id = uint_(gen) % ne;
                while (iev == id && ne != 1) id = uint_(gen) % ne;
                plan_ids.push_back(e0 + id);
            }
            for (int k = 0; k < M; k++) {
                const double random_rotation = uni(gen) * 2 * M_PI;
                double sn, cn;
                sincos(random_rotation, &sn, &cn);  // as the reference's build evaluates it
                plan_cs.push_back(cn);
                plan_cs.push_back(sn);
            }
            const int e = e0 + iev;
            rowoff[(size_t)e + 1] = rowoff[(size_t)e] + M;
            const long long n1 = off[(size_t)e + 1] - off[(size_t)e];
            for (size_t q = base; q < plan_ids.size(); q++) {
                const int p = plan_ids[q];
                r->enum_mixed += (uint64_t)(n1 * (off[(size_t)p + 1] - off[(size_t)p]));
            }
        }
        const long long ng = off[(size_t)grp_ev[g + 1]] - off[(size_t)e0];
        if (ng >= 2) r->enum_same += (uint64_t)(ng * (ng - 1) / 2);
    }
    // pair cap: counters of a fresh analysis start at 0, so it cannot bind
    // while every enumerated total stays within it
    const bool qinv = s->cfg.invariant_radius_flag == 1;
    for (uint64_t t : {r->enum_same, r->enum_mixed})
        if (t > s->D.needed + 1 || (qinv && t > s->D.needed_qinv))
            return sfail(s, HBT_ERR_CAPACITY,
                         "needed_number_of_pairs may bind (order-dependent pair cap)");
    // parts: whole groups, about part_particles input particles each
    std::vector<int> part_g{0};
    long long acc = 0;
    for (int g = 0; g < r->ngroups; g++) {
        acc += ev_off[grp_ev[g + 1]] - ev_off[grp_ev[g]];
        if (acc >= s->part_particles || g + 1 == r->ngroups) {
            part_g.push_back(g + 1);
            acc = 0;
        }
    }
    r->nparts = (int)part_g.size() - 1;
    // footprint: the histogram block and each part's block at its largest
    // (every kept particle in the fast list, one window per i tile and row)
    r->hist_foot = hist_layout(s);
    size_t largest = 0, parts_foot = 0;
    for (int k = 0; k < r->nparts; k++) {
        const int e0 = grp_ev[part_g[k]], e1 = grp_ev[part_g[k + 1]];
        size_t wins = 0;
        for (int e = e0; e < e1; e++) {
            const long long ne = off[(size_t)e + 1] - off[(size_t)e];
            wins += (size_t)((ne + s->shape.TI - 1) / s->shape.TI) *
                    (size_t)(rowoff[(size_t)e + 1] - rowoff[(size_t)e]);
        }
        const size_t f = part_layout(s, nullptr, off[(size_t)e1] - off[(size_t)e0],
                                     rowoff[(size_t)e1] - rowoff[(size_t)e0], wins,
                                     part_g[k + 1] - part_g[k]);
        largest = std::max(largest, f);
        parts_foot += f;
    }
    r->foot = r->hist_foot + parts_foot;
    Request *rq = r.get();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->hist_foot + largest > s->arena.total()) {
            s->err = "request needs more HBM than the server's arena";
            return HBT_ERR_CAPACITY;
        }
        if (s->stop) {
            s->err = "server is stopping";
            return HBT_ERR_STATE;
        }
        r->id = s->next_id++;
        *req_id = r->id;
        s->reqs[r->id] = std::move(r);
        s->st.submitted++;
    }
    // each part is queued as soon as it is prepared: its upload and launches
    // run while the next part is prepared here
    const double D = px_window(s->D);
    for (int k = 0; k < rq->nparts; k++) {
        std::unique_ptr<Part> pp(new Part());
        Part &p = *pp;
        p.req = rq;
        const int g0 = part_g[k], g1 = part_g[k + 1];
        p.e0 = grp_ev[g0];
        p.nev = grp_ev[g1] - p.e0;
        p.ngroups = g1 - g0;
        p.grp_ev.resize((size_t)p.ngroups + 1);
        for (int g = 0; g <= p.ngroups; g++) p.grp_ev[(size_t)g] = grp_ev[g0 + g] - p.e0;
        {
            std::lock_guard<std::mutex> lk(s->mu);
            if (!s->host_pool.empty()) {
                p.hp = std::move(s->host_pool.back());
                s->host_pool.pop_back();
            }
        }
        std::string err;
        int rc = gather_particles(s->D, s->cfg, p.ngroups, p.grp_ev.data(), p.nev,
                                  ev_off + p.e0, soa, p.hp, err);
        if (rc == HBT_OK) {
            p.pmax2 = p.hp.pmax2;
            const long long r0 = rowoff[(size_t)p.e0];
            p.rowoff.resize((size_t)p.nev + 1);
            for (int e = 0; e <= p.nev; e++) p.rowoff[(size_t)e] = rowoff[(size_t)(p.e0 + e)] - r0;
            const long long rows = p.rowoff[(size_t)p.nev];
            p.plan_ids.resize((size_t)rows);
            for (long long q = 0; q < rows; q++) p.plan_ids[(size_t)q] = plan_ids[(size_t)(r0 + q)] - p.e0;
            p.plan_cs.assign(plan_cs.begin() + 2 * r0, plan_cs.begin() + 2 * (r0 + rows));
            uint64_t es = 0;
            const HostParticles &hp = p.hp;
            items_same(s->shape, hp.off, hp.foff, p.grp_ev, hp.srt_cell.data(), hp.srt_py.data(),
                       D, 0, p.items, &p.item_pairs, es);
            items_mixed(s->shape, hp.foff, p.grp_ev, p.rowoff, p.plan_ids, hp.foff,
                        hp.f_pt.data(), hp.f_pt.data(), pt_window(s->D), 0, p.items, p.wins,
                        &p.item_pairs);
            for (uint64_t q : p.item_pairs) p.pairs_left += q;
        } else {
            // no work for this part; the request ends failed once its other
            // parts have run
            p.hp = HostParticles();
            p.items.clear();
            p.item_pairs.clear();
            p.wins.clear();
            p.plan_cs.clear();
            p.rowoff.assign(1, 0);
        }
        p.foot = part_layout(s, &p, p.hp.nf, (long long)p.plan_cs.size() / 2, p.wins.size(),
                             p.ngroups);
        const bool ev = hipEventCreateWithFlags(&p.done_ev, hipEventDisableTiming) == hipSuccess;
        std::lock_guard<std::mutex> lk(s->mu);
        if (rc != HBT_OK && rq->err.empty()) rq->err = err;
        if (!ev) {
            p.done_ev = nullptr;
            if (rq->err.empty()) rq->err = "hipEventCreate failed";
        }
        rq->t_queued = s->now();
        s->queue.push_back(pp.get());
        rq->parts.push_back(std::move(pp));
        s->cv.notify_all();
    }
    return HBT_OK;
}

int hbt_server_info(hbt_server *s, int64_t id, hbt_request_info *info) {
    if (!s || !info) return HBT_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    Request *r = find(s, id);
    if (!r) return HBT_ERR_ARG;
    info->state = r->state;
    info->nev = r->nev;
    info->n_kept = r->n_kept;
    info->enum_same = r->enum_same;
    info->enum_mixed = r->enum_mixed;
    info->footprint = (int64_t)r->foot;
    info->t_submit = r->t_submit;
    info->t_queued = r->t_queued;
    info->t_admit = r->t_admit;
    info->t_active = r->t_active;
    info->t_first_launch = r->t_first;
    info->t_last_launch = r->t_last;
    info->t_done = r->t_done;
    return HBT_OK;
}

int hbt_server_wait(hbt_server *s, int64_t id, double timeout_s) {
    if (!s) return HBT_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    Request *r = find(s, id);
    if (!r) return HBT_ERR_ARG;
    auto done = [&] { return r->state == HBT_REQ_DONE || r->state == HBT_REQ_FAILED; };
    if (timeout_s < 0) {
        s->cv.wait(lk, done);
    } else if (!s->cv.wait_for(lk, std::chrono::duration<double>(timeout_s), done)) {
        return HBT_ERR_STATE;
    }
    if (r->state == HBT_REQ_FAILED) {
        s->err = r->err;
        return HBT_ERR_HIP;
    }
    return HBT_OK;
}

int hbt_server_result(hbt_server *s, int64_t id, hbt_result *o) {
    if (!s || !o) return HBT_ERR_ARG;
    const char *buf = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        Request *r = find(s, id);
        if (!r) return HBT_ERR_ARG;
        if (r->state != HBT_REQ_DONE) {
            s->err = "request is not done";
            return HBT_ERR_STATE;
        }
        buf = s->h_results[r->rslot];
    }
    // the result slot stays with the request until it is released
    const double *aos = (const double *)buf;
    const unsigned long long *q = (const unsigned long long *)(buf + (size_t)s->D.nbins * 64);
    const long nq = s->D.nq;
    result_from_host(s->D, s->cfg, aos, q, q + nq, (const double *)(q + 2 * nq),
                     (const double *)(q + 3 * nq), o);
    return HBT_OK;
}

int hbt_server_release(hbt_server *s, int64_t id) {
    if (!s) return HBT_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    Request *r = find(s, id);
    if (!r) return HBT_ERR_ARG;
    if (r->state != HBT_REQ_DONE && r->state != HBT_REQ_FAILED) {
        s->err = "request is still running";
        return HBT_ERR_STATE;
    }
    if (r->rslot >= 0) s->free_results.push_back(r->rslot);
    for (auto &p : r->parts)
        if (p->done_ev) hipEventDestroy(p->done_ev);
    s->reqs.erase(id);
    s->cv.notify_all();
    return HBT_OK;
}

int hbt_server_get_stats(hbt_server *s, hbt_server_stats *out) {
    if (!s || !out) return HBT_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->st;
    out->arena_used = (int64_t)s->arena.used();
    out->arena_peak = (int64_t)s->arena.peak();
    out->queued = (int32_t)s->queue.size();  // parts
    return HBT_OK;
}

}  // extern "C"
