// The batched form: many clouds reconstructed with one launch per stage.
//
// The single-cloud pipeline (SolvePoissionReconstructionMesh,
// mesh/PoissonReconstruction.cpp:195-245) waits for the host three times:
// for the level counts, for the depth-D marching cubes' counts and at the
// end.  The batch runs the same three phases for all its clouds together.
// In each phase every cloud's stage drivers run with a Recorder installed, so
// their launches are recorded instead of queued; the recordings are then
// merged: the launches of all clouds from one launch site (and occurrence)
// become one launch of the kernel's batched form (multi_kernel, psr_launch.h),
// whose workgroups are the clouds' grids laid end to end, with every cloud's
// arguments in one device blob uploaded once per phase.  The sort of each
// cloud is thereby a segmented sort (one launch per pass for all segments),
// and the per-depth CG solves of all clouds run as independent groups of one
// persistent launch (each group stops on its own convergence), split into
// waves that fit the resident-workgroup budget.  Each cloud keeps a context
// of its own (buffers, mailbox, results) on the batch's stream; its mesh stays
// in that context's buffers until the next batch call.
//
// A phase's merged work (the blob upload and the launches) is captured once
// into a HIP graph when the same launch structure recurs, and replayed from
// then on.  The CG groups' workgroup cap is picked per batch shape by an
// online tuner (BatchTuner) that times the candidates on the first calls.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "psr_internal.h"

namespace psr {

namespace {

// Online tuner of the CG plan: for each batch shape it runs the candidate
// plans once each, in order, timing the phase that holds the solves (GPU time
// between events), then keeps the fastest.  The shape's first call runs the
// first candidate untimed.
struct BatchTuner {
    // 0: the single-cloud plan (256 rows a workgroup); w: all solves in w
    // waves of the resident budget
    static constexpr int kCandidates[] = {0, 1, 2, 4, 8};
    static constexpr int kCount = int(sizeof(kCandidates) / sizeof(int));
    struct Entry {
        float ms[kCount];
        int tried = 0;
        int best = -1;
        bool warm = false;  // the shape's first call, untimed: its CG queue order had no history
    };
    std::map<unsigned long long, Entry> table;
    static unsigned long long key(int precision, int depth, size_t clouds, size_t points) {
        auto lg = [](size_t v) {
            int b = 0;
            while (v > 1) {
                v >>= 1;
                ++b;
            }
            return b;
        };
        return (unsigned long long)precision << 40 | (unsigned long long)depth << 32 |
               (unsigned long long)lg(clouds) << 16 | (unsigned long long)lg(points);
    }
    // The candidate index to run for this shape (exploring first).
    int pick(unsigned long long k) {
        Entry& e = table[k];
        return e.best >= 0 ? e.best : e.tried;
    }
    void report(unsigned long long k, int idx, float ms) {
        Entry& e = table[k];
        if (!e.warm) {
            e.warm = true;
            return;
        }
        if (e.best >= 0 || idx != e.tried || !(ms > 0.f)) return;
        e.ms[idx] = ms;
        if (++e.tried == kCount) e.best = int(std::min_element(e.ms, e.ms + kCount) - e.ms);
    }
};
constexpr int BatchTuner::kCandidates[];

BatchTuner& tuner() {
    static BatchTuner t;
    return t;
}

}  // namespace

class Batch {
  public:
    explicit Batch(psr_ctx* owner) : owner_(owner) {
        // a fixed CG plan instead of the tuner's (0: the single-cloud plan)
        if (owner->c.cfg.batch_cg_plan >= 0) fixedPlan_ = owner->c.cfg.batch_cg_plan;
    }
    psr_ctx* cloud(uint32_t i) const { return i < clouds_.size() ? clouds_[i] : nullptr; }
    ~Batch() {
        for (auto& g : graphs_)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
        for (psr_ctx* c : clouds_) psr_destroy(c);
        if (hostBlob_) (void)hipHostFree(hostBlob_);
    }
    int run(const psr_points* pts, uint32_t count, psr_mesh* outs, std::vector<psr_stats>& stats,
            std::string& err);
    int lastPlan() const { return lastPlan_; }

  private:
    struct Merged {
        const void* fn;
        unsigned grid, block, shmem;
        size_t packOff, counterOff;
        MultiStarts starts;
        bool co;
    };
    struct Graph {
        std::vector<unsigned long long> sig;
        hipGraphExec_t exec = nullptr;
        int seen = 0;
    };
    psr_ctx* owner_;
    std::vector<psr_ctx*> clouds_;
    char* hostBlob_ = nullptr;
    size_t hostBlobBytes_ = 0;
    std::map<std::pair<const void*, unsigned>, int> budget_;
    std::vector<Graph> graphs_;
    int lastPlan_ = 0;
    int fixedPlan_ = -1;

    Ctx& bc() { return owner_->c; }
    int resident(const void* fn, unsigned block, unsigned shmem);
    void flush(Recorder& rec, const std::vector<int>& slot, bool timeCg);
    void run_merged(const std::vector<Merged>& m, const char* blob, size_t bytes, bool timeCg);
};

int Batch::resident(const void* fn, unsigned block, unsigned shmem) {
    auto k = std::make_pair(fn, shmem);
    auto it = budget_.find(k);
    if (it != budget_.end()) return it->second;
    int per = 0, cus = 0;
    PSR_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, fn, int(block), shmem));
    PSR_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, bc().device));
    return budget_[k] = std::max(1, per * cus);
}

// Merges the clouds' recordings (slot[k]: the cloud of recording k) into
// launches of the batched kernels and queues them.
void Batch::flush(Recorder& rec, const std::vector<int>& slot, bool timeCg) {
    // The CG's queue takes the clouds' groups in launch order, and the launch
    // lasts as long as the group that starts last plus its solve: the clouds
    // whose previous solve in this context took the most iterations go first
    // (longest first; the same order while the batch repeats).  A context's
    // first call has no history and keeps the batch order.
    std::vector<long long> cgWeight(rec.clouds.size(), 0);
    for (size_t k = 0; k < rec.clouds.size() && k < slot.size(); ++k) {
        const Ctx& c = clouds_[size_t(slot[k])]->c;
        if (c.runs == 0) continue;
        for (int d = 0; d <= std::min(c.cfg.depth, PSR_MAX_DEPTH); ++d)
            cgWeight[k] = std::max<long long>(cgWeight[k], c.hostResults->cgIters[d]);
    }
    // launch order: every cloud's launches in its own order; a launch site
    // (and occurrence) a cloud skipped is placed right after that cloud's
    // previous launch
    std::vector<std::pair<const void*, int>> order;
    for (const auto& v : rec.clouds) {
        int prev = -1;
        for (const LaunchRec& r : v) {
            const auto key = std::make_pair(r.site, r.occurrence);
            int pos = -1;
            for (int i = 0; i < int(order.size()); ++i)
                if (order[i] == key) {
                    pos = i;
                    break;
                }
            if (pos < 0) {
                pos = prev + 1;
                order.insert(order.begin() + pos, key);
            } else if (pos <= prev) {
                throw std::logic_error("batch clouds recorded launches in different orders");
            }
            prev = pos;
        }
    }
    std::vector<char> blob;
    auto align = [&](size_t a) { blob.resize((blob.size() + a - 1) / a * a); };
    std::vector<Merged> merged;
    std::vector<const LaunchRec*> items;
    std::vector<std::pair<long long, const LaunchRec*>> weighted;
    for (const auto& key : order) {
        items.clear();
        weighted.clear();
        for (size_t k = 0; k < rec.clouds.size(); ++k)
            for (const LaunchRec& r : rec.clouds[k])
                if (r.site == key.first && r.occurrence == key.second && r.grid > 0) weighted.emplace_back(cgWeight[k], &r);
        if (weighted.empty()) continue;
        if (weighted[0].second->coResident)
            std::stable_sort(weighted.begin(), weighted.end(),
                             [](const auto& a, const auto& b) { return a.first > b.first; });
        for (const auto& w : weighted) items.push_back(w.second);
        const LaunchRec& r0 = *items[0];
        unsigned shmem = 0;
        for (const LaunchRec* r : items) {
            if (r->multiFn != r0.multiFn || r->block != r0.block || r->pack.size() != r0.pack.size())
                throw std::logic_error("batch clouds recorded different kernels at one launch site");
            shmem = std::max(shmem, r->shmem);
        }
        // at most kMaxMultiClouds clouds a launch; a co-resident launch (the
        // CG) runs as a persistent queue of at most the resident budget
        const long long cap = r0.coResident ? resident(r0.multiFn, r0.block, shmem) : (1ll << 31) - 1;
        for (size_t i0 = 0; i0 < items.size();) {
            size_t i1 = i0;
            long long wgs = 0;
            while (i1 < items.size() && i1 - i0 < size_t(kMaxMultiClouds)) wgs += items[i1++]->grid;
            for (size_t i = i0; i < i1; ++i)
                if (r0.coResident && items[i]->grid > cap)
                    throw std::runtime_error("a cloud's CG exceeds the resident workgroup budget");
            Merged m{};
            m.fn = r0.multiFn;
            m.grid = r0.coResident ? unsigned(std::min<long long>(wgs, cap)) : unsigned(wgs);
            m.block = r0.block;
            m.shmem = shmem;
            m.co = r0.coResident;
            m.starts.n = int(i1 - i0);
            unsigned acc = 0;
            for (size_t i = i0; i <= i1; ++i) {
                m.starts.s[i - i0] = acc;
                if (i < i1) acc += items[i]->grid;
            }
            align(256);
            m.packOff = blob.size();
            for (size_t i = i0; i < i1; ++i) blob.insert(blob.end(), items[i]->pack.begin(), items[i]->pack.end());
            if (m.co) {  // the queue's counter, zero in the uploaded blob
                align(256);
                m.counterOff = blob.size();
                blob.resize(blob.size() + sizeof(unsigned), 0);
            }
            merged.push_back(m);
            i0 = i1;
        }
    }
    if (merged.empty()) return;
    run_merged(merged, blob.data(), blob.size(), timeCg);
}

// Uploads the argument blob and queues the merged launches; a launch
// structure seen before runs from its captured graph.
void Batch::run_merged(const std::vector<Merged>& m, const char* blob, size_t bytes, bool timeCg) {
    Ctx& c = bc();
    hipStream_t s = c.stream;
    if (bytes > hostBlobBytes_) {
        if (hostBlob_) PSR_HIP_CHECK(hipHostFree(hostBlob_));
        hostBlob_ = nullptr;
        hostBlobBytes_ = 0;
        PSR_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&hostBlob_), bytes + bytes / 2 + 4096, hipHostMallocDefault));
        hostBlobBytes_ = bytes + bytes / 2 + 4096;
    }
    // the device blob keeps its address while the structure repeats (a
    // graph's nodes hold it); it only grows
    char* dev = c.buf<char>("batch_blob", hostBlobBytes_);
    std::memcpy(hostBlob_, blob, bytes);
    bool anyCo = false;
    for (const Merged& x : m) anyCo |= x.co;
    // the structure: kernels, grids, shared memory, offsets and the blob's addresses
    std::vector<unsigned long long> sig;
    sig.push_back(reinterpret_cast<unsigned long long>(dev));
    sig.push_back(reinterpret_cast<unsigned long long>(hostBlob_));
    sig.push_back(bytes);
    for (const Merged& x : m) {
        sig.push_back(reinterpret_cast<unsigned long long>(x.fn));
        sig.push_back((unsigned long long)x.grid << 32 | x.shmem);
        sig.push_back((unsigned long long)x.packOff << 1 | (x.co ? 1u : 0u));
        sig.push_back((unsigned long long)x.counterOff);
        sig.push_back((unsigned long long)x.starts.n);
        for (int i = 0; i <= x.starts.n; ++i) sig.push_back(x.starts.s[i]);
    }
    if (anyCo) cg_admit_device(c.ev[9]);
    auto queue = [&]() {
        PSR_HIP_CHECK(hipMemcpyAsync(dev, hostBlob_, bytes, hipMemcpyHostToDevice, s));
        for (const Merged& x : m) {
            const void* packs = dev + x.packOff;
            MultiStarts starts = x.starts;
            unsigned* counter = reinterpret_cast<unsigned*>(dev + x.counterOff);
            void* args[] = {&packs, &starts, &counter};  // the queue form (x.co) takes the counter too
            void* args2[] = {&packs, &starts};
            PSR_HIP_CHECK(hipLaunchKernel(x.fn, dim3(x.grid), dim3(x.block), x.co ? args : args2, x.shmem, s));
        }
    };
    Graph* g = nullptr;
    for (Graph& q : graphs_)
        if (q.sig == sig) g = &q;
    if (!g) {
        if (graphs_.size() >= 16) {  // keep the most recent structures
            if (graphs_.front().exec) PSR_HIP_CHECK(hipGraphExecDestroy(graphs_.front().exec));
            graphs_.erase(graphs_.begin());
        }
        graphs_.push_back(Graph{sig, nullptr, 0});
        g = &graphs_.back();
    }
    ++g->seen;
    if (timeCg) PSR_HIP_CHECK(hipEventRecord(c.ev[10], s));
    if (g->exec) {
        PSR_HIP_CHECK(hipGraphLaunch(g->exec, s));
    } else if (g->seen >= 2) {
        hipGraph_t graph = nullptr;
        PSR_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
            queue();
        } catch (...) {
            (void)hipStreamEndCapture(s, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        PSR_HIP_CHECK(hipStreamEndCapture(s, &graph));
        PSR_HIP_CHECK(hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0));
        PSR_HIP_CHECK(hipGraphDestroy(graph));
        PSR_HIP_CHECK(hipGraphLaunch(g->exec, s));
    } else {
        queue();
    }
    if (timeCg) PSR_HIP_CHECK(hipEventRecord(c.ev[11], s));
    if (anyCo) {
        PSR_HIP_CHECK(hipEventRecord(c.ev[9], s));
        cg_hold_device(c.ev[9]);
    }
}

int Batch::run(const psr_points* pts, uint32_t count, psr_mesh* outs, std::vector<psr_stats>& stats,
               std::string& err) {
    Ctx& b = bc();
    while (clouds_.size() < count) {
        psr_ctx* w = nullptr;
        int rc = psr_create(&b.cfg, b.stream, &w);
        if (rc != PSR_OK) return err = "batch cloud context creation failed", rc;
        clouds_.push_back(w);
    }
    stats.assign(count, psr_stats{});
    int status = PSR_OK;
    auto record_error = [&](uint32_t i, int rc, const std::string& m) {
        if (status == PSR_OK) {
            status = rc;
            err = "cloud " + std::to_string(i) + ": " + m;
        }
    };
    for (uint32_t i = 0; i < count; ++i) outs[i] = psr_mesh{nullptr, nullptr, 0, 0};
    const uint32_t pass = b.cfg.batch_workers ? b.cfg.batch_workers : 64;
    for (uint32_t i0 = 0; i0 < count; i0 += pass) {
        const uint32_t i1 = std::min(count, i0 + pass);
        std::vector<uint32_t> live;
        size_t points = 0;
        for (uint32_t i = i0; i < i1; ++i) {
            psr_ctx* w = clouds_[i];
            const int rc = call_begin(w->c, pts[i], w->lastError);
            if (rc != PSR_OK) {
                record_error(i, rc, w->lastError);
                continue;
            }
            live.push_back(i);
            points += pts[i].count;
        }
        if (live.empty()) continue;
        // the tuner's cap of the CG groups for this batch shape
        const unsigned long long tkey = BatchTuner::key(b.cfg.precision, b.cfg.depth, live.size(), points);
        const int cand = fixedPlan_ >= 0 ? -1 : tuner().pick(tkey);
        lastPlan_ = cand < 0 ? fixedPlan_ : BatchTuner::kCandidates[cand];
        Recorder rec;
        Recorder*& cur = current_recorder();
        // one phase: every live cloud's drivers recorded, then merged and queued
        auto phase = [&](auto&& drive, bool timeCg) {  // drive(ctx, cloud index)
            rec.clouds.assign(live.size(), {});
            std::vector<int> slot(live.size());
            std::vector<uint32_t> keep;
            std::vector<uint32_t> immediate;
            cur = &rec;
            try {
                for (size_t k = 0; k < live.size(); ++k) {
                    rec.cloud = int(k);
                    slot[k] = int(live[k]);
                    Ctx& c = clouds_[live[k]]->c;
                    c.arena.defer_frees(true);
                    try {
                        drive(c, live[k]);
                        keep.push_back(live[k]);
                    } catch (const NotRecordable&) {
                        rec.clouds[k].clear();
                        immediate.push_back(live[k]);
                    } catch (const std::invalid_argument& e) {
                        rec.clouds[k].clear();
                        record_error(live[k], PSR_ERR_INVALID, e.what());
                    } catch (const HipError&) {
                        throw;
                    } catch (const std::exception& e) {
                        rec.clouds[k].clear();
                        record_error(live[k], PSR_ERR_INTERNAL, e.what());
                    }
                }
                cur = nullptr;
            } catch (...) {
                cur = nullptr;
                for (uint32_t i : live) clouds_[i]->c.arena.defer_frees(false);
                throw;
            }
            flush(rec, slot, timeCg);
            for (uint32_t i : live) clouds_[i]->c.arena.defer_frees(false);
            // clouds whose stage cannot be recorded (a subdivision in several
            // working sets waits for the host between them) run on their own
            for (uint32_t i : immediate) {
                try {
                    drive(clouds_[i]->c, i);
                    keep.push_back(i);
                } catch (const HipError&) {
                    throw;
                } catch (const std::exception& e) {
                    record_error(i, PSR_ERR_INTERNAL, e.what());
                }
            }
            std::sort(keep.begin(), keep.end());
            live = keep;
            b.wait();
        };
        phase([&](Ctx& c, uint32_t i) { build_octree_a(c, pts[i]); }, false);
        // the CG plan: w waves of the several-rows kernel's resident budget
        // hold every depth of every cloud, each depth's rows split evenly
        int target = 0;
        if (lastPlan_ > 0 && !live.empty()) {
            size_t rows = 0, groups = 0;
            for (uint32_t i : live) {
                const Ctx& c = clouds_[i]->c;
                // nodes: the root, and 8 per occupied node of the level above
                // (mailbox slot d: occupied nodes at level d, after the level counts' wait)
                rows += 1;
                for (int d = 0; d < c.cfg.depth; ++d) rows += size_t(8) * size_t(d == 0 ? 1 : c.hostPinned[d]);
                groups += size_t(c.cfg.depth) + 1;
            }
            const long long slots = (long long)lastPlan_ * cg_batch_budget(b) - (long long)groups;
            target = int(std::min<size_t>(size_t(1) << 30, (rows + size_t(std::max(1ll, slots)) - 1) /
                                                               size_t(std::max(1ll, slots))));
            target = std::max(64, (target + 63) / 64 * 64);
        }
        for (uint32_t i : live) clouds_[i]->c.cgRowTarget = target;
        phase(
            [&](Ctx& c, uint32_t) {
                build_octree_b(c);
                compute_field(c);
                solve_depths(c);
                points_implicit(c);
                build_geometry(c);
                extract_mesh_a(c);
            },
            true);
        if (cand >= 0) tuner().report(tkey, cand, elapsed(b.ev[10], b.ev[11]));
        phase([&](Ctx& c, uint32_t) { extract_mesh_b(c); }, false);
        for (uint32_t i : live) {
            psr_ctx* w = clouds_[i];
            const int rc = call_end(w->c, &outs[i], false, w->lastError);
            if (rc != PSR_OK) {
                outs[i] = psr_mesh{nullptr, nullptr, 0, 0};
                record_error(i, rc, w->lastError);
                continue;
            }
            stats[i] = w->c.stats;
        }
    }
    return status;
}

Batch* batch_create(psr_ctx* owner) { return new Batch(owner); }
void batch_destroy(Batch* b) { delete b; }
int batch_run(Batch* b, const psr_points* pts, uint32_t count, psr_mesh* outs, std::vector<psr_stats>& stats,
              std::string& err) {
    return b->run(pts, count, outs, stats, err);
}
int batch_last_plan(const Batch* b) { return b ? b->lastPlan() : 0; }
psr_ctx* batch_cloud(const Batch* b, uint32_t i) { return b ? b->cloud(i) : nullptr; }

}  // namespace psr
