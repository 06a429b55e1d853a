// cvt_internal.h -- library-internal helpers shared by the C ABI translation
// units (not part of the public header).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "cvt.h"

namespace cvt {

// Records the calling thread's error message (cvt_last_error) and returns code.
__attribute__((visibility("hidden"))) int fail(int code, const char* fmt, ...);

// Device memory arena of a serving loop: chunks taken from hipMalloc as it
// grows, carved first-fit (256-byte aligned) and returned by free() with
// neighbours coalesced inside a chunk.  Jobs created and retired continuously
// then never call hipMalloc / hipFree in the steady state (hipFree may wait for
// the whole device, stalling every other job's work).  Thread-safe.
struct DevArena {
    struct Chunk {
        uintptr_t base;
        size_t size;
    };
    std::mutex mu;
    std::vector<Chunk> chunks;
    std::map<uintptr_t, size_t> free_;  // start -> length (never spanning two chunks)
    size_t grow_bytes = size_t(1) << 31;
    size_t reserved = 0;  // bytes held from hipMalloc
    void* alloc(size_t bytes);
    void free(void* p, size_t bytes);
    ~DevArena();
};

// The arena the calling thread allocates new jobs, LBVHs, batches and
// temporaries from (nullptr: plain hipMalloc).  Objects remember the arena
// they were created with and return their memory to it.
__attribute__((visibility("hidden"))) DevArena*& current_arena();
__attribute__((visibility("hidden"))) int dev_alloc(DevArena* a, void** p, size_t bytes);
__attribute__((visibility("hidden"))) void dev_free(DevArena* a, void* p, size_t bytes);

// Device temporaries handed out in call order.  reset() starts a new round of
// get() calls that reuse the previous round's buffers (grown when too small);
// the owner must have synchronised all work that used them.  A pool that is
// not reused simply frees everything when destroyed.
struct TempPool {
    struct Slot {
        void* p = nullptr;
        size_t cap = 0;
    };
    std::vector<Slot> slots;
    size_t next = 0;
    DevArena* arena = current_arena();
    TempPool() = default;
    // one owner per buffer: a copy would free the buffers the other still uses
    TempPool(const TempPool&) = delete;
    TempPool& operator=(const TempPool&) = delete;
    TempPool(TempPool&& o) noexcept : slots(std::move(o.slots)), next(o.next), arena(o.arena) {
        o.slots.clear();
        o.next = 0;
    }
    TempPool& operator=(TempPool&&) = delete;
    void reset() { next = 0; }
    void* get(size_t bytes) {
        bytes = std::max<size_t>(bytes, 256);
        if (next == slots.size()) slots.emplace_back();
        Slot& s = slots[next++];
        if (s.cap < bytes) {
            if (s.p) dev_free(arena, s.p, s.cap);
            const size_t want = std::max(bytes, s.cap + s.cap / 2);
            s.p = nullptr;
            s.cap = 0;
            if (dev_alloc(arena, &s.p, want)) return nullptr;
            s.cap = want;
        }
        return s.p;
    }
    size_t bytes() const {
        size_t b = 0;
        for (const Slot& s : slots) b += s.cap;
        return b;
    }
    ~TempPool() {
        for (Slot& s : slots)
            if (s.p) dev_free(arena, s.p, s.cap);
    }
};

// One-sided Hausdorff of the remeshed surface with caller-owned temporaries
// (the serving loop keeps one pool per check worker).  tol >= 0 with `within`:
// first a bounded query decides whether every probe lies within tol
// (*within); the exact distance (h_out, nullable) is then computed only when
// it does or exact_always is set (else *h_out = -1).  tol < 0: exact only.
__attribute__((visibility("hidden"))) int job_hausdorff(cvt_job* j, double eps, TempPool& pool, double tol,
                                                        int* within, bool exact_always, double* h_out, int* nF_out);
// The fast convergence verdict of job_hausdorff, enqueued on the job's stream
// with pool-held temporaries: *enqueued = 1, and *flag_host receives 0 (every
// probe within tol), 1 (not) or 2 (undecided) once that stream is
// synchronised (the pool must not be reused before then); *enqueued = 0 when
// the fast path does not apply (no tolerance, or nS >= 2^21).
// keys_out/site_n_out (nullable) receive its remesh-cell keys and site normals
// (pool buffers), which a remesh of the same sites can reuse.
__attribute__((visibility("hidden"))) int check_fast_enqueue(cvt_job* j, double eps, TempPool& pool, double tol,
                                                             int* flag_host, int* enqueued,
                                                             unsigned long long** keys_out = nullptr,
                                                             float4** site_n_out = nullptr);

// A batch of jobs for a serving loop, over job slots and block tables of
// fixed capacity: jobs join and leave by editing the device slot and block
// tables in place (a joining job's descriptor and block references are
// written, a leaving job's reset to empty), and the captured iteration graph
// is kept, with the grids of its table-spanning kernel nodes set to the
// extent of the tables in use (hipGraphExecKernelNodeSetParams), so a steady
// stream of jobs runs without capturing the graph again and without
// launching blocks past the last job.  The graph is captured again only when
// a table outgrows its capacity (all then grow to four times what they hold);
// a joining job that finds no contiguous room, or a site table whose extent
// is over twice what it holds, is compacted in place.  Jobs must share K and
// kproj, must not run elsewhere while in the batch, and are added and removed
// only while no step of the batch is in flight.
struct SlotBatch;
struct SlotBatchCounters {
    long captures, first, grow;  // graph captures: total, first, for growth
    long compactions;            // tables laid out again in place
    long extent_updates;         // grid edits of the instantiated graph
};
// First capacities: jobs0 jobs, site_blocks0 blocks of 256 sites (0: defaults).
__attribute__((visibility("hidden"))) int slot_batch_create(int K, int kproj, int jobs0, int site_blocks0,
                                                            SlotBatch** out);
__attribute__((visibility("hidden"))) int slot_batch_add(SlotBatch* b, cvt_job* job);
__attribute__((visibility("hidden"))) int slot_batch_remove(SlotBatch* b, cvt_job* job);
// Runs iters iterations of every job in the batch and waits for them.
__attribute__((visibility("hidden"))) int slot_batch_step(SlotBatch* b, int iters);
__attribute__((visibility("hidden"))) SlotBatchCounters slot_batch_counters(const SlotBatch* b);
__attribute__((visibility("hidden"))) int slot_batch_size(const SlotBatch* b);
__attribute__((visibility("hidden"))) void slot_batch_destroy(SlotBatch* b);

}  // namespace cvt
