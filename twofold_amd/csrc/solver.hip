// solver.hip - gfx950 kernels and host driver of the soft-step contact solver, behind the C ABI of
// include/twofold_solver.h.
//
// One tfs_step runs src/world.cpp:1591-1691 and updateSleep (1174-1247) on state that stays in HBM:
//
//   step state, inverse mass (per body)            k_step_init
//   active / fast / fast-solid lists (per contact)  k_classify + stream compaction
//   prepare + impact marking (per active contact)   k_prepare, k_impact_flag + compaction
//   greedy colouring in active order                k_ep_* (per-body endpoint ranks) + k_color
//   row order: colour, then point count             k_row_hist + scan + k_row_scatter
//   row packing (SoA rows, coalesced)               k_pack
//   per substep: [refresh] integrate velocities, warm start, biased solve (+impact), integrate
//   positions, relax (+impact) - one launch per colour; contacts outside the colours run in order
//   restitution, final speed limit, impulses back to the contacts, sleep and islands.
//
// Bodies live in three device arrays: the full tfs_body records (cold fields), a 64-byte motion
// record (v, w, dp, dq) that the row kernels gather and scatter, and a 96-byte mass record (world
// inverse mass and inertia) that they gather. Coloured contacts are packed once per step into
// structure-of-arrays rows in row order, so every pass reads them with coalesced 16-byte loads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "stages.h"
#include "collide3d.h"
#include "joints.h"

using namespace tf;

namespace {

constexpr int kThreads = 256;
constexpr int kScanSlots = 16;         // scans in one step
constexpr int kProfileFamilies = 8;
constexpr int kKeys = 5 * TFS_COLORS + 1; // row-order buckets: colour x point count, then the rest
constexpr uint32_t kSpinLimit = 1u << 26;
constexpr int kCollideThreads = 64;    // the device narrowphase: one wave per work-group
constexpr uint32_t kOrderMax = 16384;  // longest fast list put in collision order
constexpr int kIslandRounds = 64;      // hook-and-jump rounds of the island pass before it gives up
constexpr int kIslandAhead = 6;        // rounds launched before the host checks the first time (at most)
// The per-body kernels of a substep run a wave per block: a 32k-body pile is then 513 blocks over the
// 256 compute units rather than 129 blocks of four waves on half of them.
constexpr uint32_t kBodyThreads = 64;

struct DevCounts {
    uint32_t nActive, nFast, nFastSolid, nImpact;
    uint32_t pointsActive, pointsImpact;
    uint32_t scanTotal;
    uint32_t error;
    uint32_t anyRestitution;
    uint32_t ticket;
    uint32_t restTicket[2];   // level dataflow of the active and impact rest lists
    uint32_t restLevels[2];
    uint32_t restNarrow[2];   // ~(first level no wider than a rest work-group), 0 when every level is wider
    uint32_t restShape[2][4]; // per rest list: planar contacts, their points, 3D contacts, their points
    uint32_t scanTicket[kScanSlots];
    uint32_t changed[kIslandRounds];
    uint32_t nTouchEvents, nHits, touching;   // reportTouches of the step
    uint32_t streamPairs, streamRest;         // contacts the stream's pairs and resting contacts leave
    uint32_t nJointsActive, nJointsBegun;     // active joints now (after breaking), and at the step's start
    uint32_t colorStart[TFS_COLORS + 1];
    uint32_t impactColorStart[TFS_COLORS + 1];
    uint32_t colorPoints[TFS_COLORS + 1];
    uint32_t impactColorPoints[TFS_COLORS + 1];
};

// ---- body records ----------------------------------------------------------------------------

__device__ __forceinline__ Motion loadMotion(const float4* m, uint32_t i) {
    const float4 v = m[4 * i], w = m[4 * i + 1], dp = m[4 * i + 2], dq = m[4 * i + 3];
    return Motion{ vec(v.x, v.y, v.z), vec(w.x, w.y, w.z), vec(dp.x, dp.y, dp.z), quat(dq.x, dq.y, dq.z, dq.w) };
}
__device__ __forceinline__ void storeVelocity(float4* m, uint32_t i, const Motion& x) {
    m[4 * i] = make_float4(x.v.x, x.v.y, x.v.z, 0.0f);
    m[4 * i + 1] = make_float4(x.w.x, x.w.y, x.w.z, 0.0f);
}
__device__ __forceinline__ void storeMotion(float4* m, uint32_t i, const Motion& x) {
    storeVelocity(m, i, x);
    m[4 * i + 2] = make_float4(x.dp.x, x.dp.y, x.dp.z, 0.0f);
    m[4 * i + 3] = make_float4(x.dq.x, x.dq.y, x.dq.z, x.dq.w);
}
// 24 floats: world inverse mass (9), world inverse inertia (9), inverse mass, plane inverse mass,
// plane inverse inertia, 3 spare.
__device__ __forceinline__ Mass loadPlanarMass(const float4* m, uint32_t i) {
    const float4 e = m[6 * i + 4], f = m[6 * i + 5];
    Mass x{};
    x.invMass = e.z;
    x.pim = e.w;
    x.pii = f.x;
    return x;
}
__device__ __forceinline__ Mass loadMass(const float4* m, uint32_t i) {
    const float4 a = m[6 * i], b = m[6 * i + 1], c = m[6 * i + 2], d = m[6 * i + 3], e = m[6 * i + 4], f = m[6 * i + 5];
    Mass x;
    x.imw = M3{ { vec(a.x, a.y, a.z), vec(a.w, b.x, b.y), vec(b.z, b.w, c.x) } };
    x.iiw = M3{ { vec(c.y, c.z, c.w), vec(d.x, d.y, d.z), vec(d.w, e.x, e.y) } };
    x.invMass = e.z;
    x.pim = e.w;
    x.pii = f.x;
    return x;
}
__device__ __forceinline__ void storeMass(float4* m, uint32_t i, const tfs_body& b) {
    const float* w = b.inv_mass_world;
    const float* n = b.inv_inertia_world;
    m[6 * i] = make_float4(w[0], w[1], w[2], w[3]);
    m[6 * i + 1] = make_float4(w[4], w[5], w[6], w[7]);
    m[6 * i + 2] = make_float4(w[8], n[0], n[1], n[2]);
    m[6 * i + 3] = make_float4(n[3], n[4], n[5], n[6]);
    m[6 * i + 4] = make_float4(n[7], n[8], b.inv_mass, b.plane_inv_mass);
    m[6 * i + 5] = make_float4(b.plane_inv_inertia, 0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ uint32_t gid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ uint32_t gstride() { return gridDim.x * blockDim.x; }

// Adds every lane's v to *p with one atomic per wave (a counter that every thread of a launch adds to
// would otherwise serialize them). Called by all lanes of the wave.
__device__ __forceinline__ void waveAdd(uint32_t* p, uint32_t v) {
    for (int off = 32; off > 0; off >>= 1) { v += (uint32_t)__shfl_xor((int)v, off); }
    if ((threadIdx.x % 64) == 0 && v != 0) { atomicAdd(p, v); }
}

// ---- scans and compaction ---------------------------------------------------------------------
//
// One launch per scan, decoupled look-back: each block takes a tile of 4096 items in ticket order,
// scans it in LDS, publishes its tile total, then adds up its predecessors' published totals (or
// the first inclusive prefix it meets) and publishes its own inclusive prefix. A tile only waits on
// tiles that took earlier tickets, which publish their totals without waiting on anything, so every
// wait ends. Status words carry the launch's epoch, so old words never need clearing.

__device__ __forceinline__ uint32_t countOf(const uint32_t* nPtr, uint32_t n) { return nPtr ? *nPtr : n; }

constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr uint32_t kScanTile = kScanThreads * kScanItems;
constexpr uint32_t kFlagTotal = 1, kFlagPrefix = 2;
constexpr uint32_t kEpochLimit = 1u << 30;

struct ScanJob {
    const uint32_t* in;
    const uint32_t* nPtr;  // item count on the device, or null for nConst
    uint32_t nConst;
    uint32_t* out;   // exclusive prefix (may be `in`), or null
    uint32_t* list;  // i written to list[prefix] where in[i] != 0, or null
    uint32_t* total;
};
struct ScanJobs {
    ScanJob job[3];
    uint32_t tiles;  // tiles per job
};

__device__ __forceinline__ unsigned long long statusWord(uint32_t epoch, uint32_t flag, uint32_t value) {
    return ((unsigned long long)((epoch << 2) | flag) << 32) | value;
}

// LDS index of item i of a tile, padded so that a thread's 16 consecutive items fall in different banks.
__device__ __forceinline__ uint32_t padded(uint32_t i) { return i + i / kScanItems; }

__global__ void __launch_bounds__(kScanThreads) k_scan(ScanJobs jobs, unsigned long long* status, uint32_t epoch,
                                                       uint32_t* ticket, DevCounts* cnt) {
    __shared__ uint32_t lds[kScanTile + kScanTile / kScanItems];
    __shared__ uint32_t waveSum[kScanThreads / 64];
    __shared__ uint32_t order, tileOffset;
    const uint32_t t = threadIdx.x, lane = t % 64, wave = t / 64;
    if (t == 0) { order = atomicAdd(ticket, 1u); }
    __syncthreads();
    const uint32_t jobNo = order / jobs.tiles, tile = order % jobs.tiles;
    const ScanJob J = jobs.job[jobNo];
    const uint32_t n = countOf(J.nPtr, J.nConst);
    const uint32_t base = tile * kScanTile;
    unsigned long long* st = status + (size_t)jobNo * jobs.tiles;

    uint32_t v[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t i = base + (uint32_t)k * kScanThreads + t;
        v[k] = i < n ? J.in[i] : 0u;
        lds[padded((uint32_t)k * kScanThreads + t)] = v[k];
    }
    __syncthreads();
    // each thread's 16 consecutive items: exclusive prefixes in place, and their sum
    uint32_t run = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint32_t at = padded(t * kScanItems + (uint32_t)j);
        const uint32_t x = lds[at];
        lds[at] = run;
        run += x;
    }
    uint32_t incl = run;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)incl, off);
        if (lane >= off) { incl += y; }
    }
    if (lane == 63) { waveSum[wave] = incl; }
    __syncthreads();
    uint32_t before = 0, tileTotal = 0;
    for (uint32_t w = 0; w < kScanThreads / 64; ++w) {
        if (w < wave) { before += waveSum[w]; }
        tileTotal += waveSum[w];
    }
    const uint32_t threadBefore = before + incl - run;

    if (wave == 0) {
        uint32_t offset = 0;
        if (tile == 0) {
            if (lane == 0) { __hip_atomic_store(&st[0], statusWord(epoch, kFlagPrefix, tileTotal), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
        } else {
            if (lane == 0) { __hip_atomic_store(&st[tile], statusWord(epoch, kFlagTotal, tileTotal), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
            int look = (int)tile - 1;  // lane l reads tile look - l
            uint32_t spins = 0;
            bool done = false;
            while (!done) {
                const int j = look - (int)lane;
                uint32_t flag = kFlagPrefix, value = 0;  // before tile 0: an empty prefix
                if (j >= 0) {
                    const unsigned long long w = __hip_atomic_load(&st[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const uint32_t hi = (uint32_t)(w >> 32);
                    flag = (hi >> 2) == epoch ? (hi & 3u) : 0u;
                    value = (uint32_t)w;
                }
                if (__ballot(flag == 0u) != 0ull) {
                    if (++spins > kSpinLimit) {
                        if (lane == 0) { atomicOr(&cnt->error, 4u); }
                        done = true;
                    } else {
                        __builtin_amdgcn_s_sleep(1);
                    }
                    continue;
                }
                const unsigned long long prefixes = __ballot(flag == kFlagPrefix);
                const int stop = prefixes ? __ffsll((long long)prefixes) - 1 : 63;
                uint32_t add = (int)lane <= stop ? value : 0u;
                for (int off = 32; off > 0; off >>= 1) { add += (uint32_t)__shfl_xor((int)add, off); }
                offset += add;
                if (prefixes) {
                    done = true;
                } else {
                    look -= 64;
                }
            }
            if (lane == 0) {
                __hip_atomic_store(&st[tile], statusWord(epoch, kFlagPrefix, offset + tileTotal), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        if (lane == 0) { tileOffset = offset; }
    }
    __syncthreads();
    const uint32_t add = tileOffset + threadBefore;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) { lds[padded(t * kScanItems + (uint32_t)j)] += add; }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint32_t i = base + (uint32_t)k * kScanThreads + t;
        if (i < n) {
            const uint32_t pre = lds[padded((uint32_t)k * kScanThreads + t)];
            if (J.out) { J.out[i] = pre; }
            if (J.list && v[k]) { J.list[pre] = i; }
        }
    }
    if (t == 0 && tile == jobs.tiles - 1 && J.total) { *J.total = tileOffset + tileTotal; }
}

// ---- step state and lists ------------------------------------------------------------------------

// src/world.cpp:1591-1601: pose at step start, travel bound, dp and dq reset, inverse mass.
// Also clears the per-body scratch of the step (impact marks, colouring words and degrees, warm-start
// degrees and cursors), which saves a memset each.
struct BodyScratch {
    uint8_t* struck;
    uint32_t *deg, *cursor, *bdeg, *bcursor;
    unsigned long long* word;
};

__global__ void k_step_init(tfs_body* bodies, uint32_t n, float dt, float4* motion, float4* mass, uint8_t* moving,
                            BodyScratch z) {
    for (uint32_t i = gid(); i < n; i += gstride()) {
        z.struck[i] = 0;
        z.deg[i] = z.cursor[i] = z.bdeg[i] = z.bcursor[i] = 0;
        z.word[i] = 0ull;
        tfs_body b = bodies[i];
        if (b.alive) {
            for (int k = 0; k < 4; ++k) {
                b.q0[k] = b.q[k];
                b.q_base[k] = b.q[k];
            }
            for (int k = 0; k < 3; ++k) { b.com0[k] = b.com[k]; }
            b.expected_travel = (length(ld3(b.v)) + length(ld3(b.w)) * b.extent) * dt;
            b.dp[0] = b.dp[1] = b.dp[2] = 0.0f;
            b.dq[0] = b.dq[1] = b.dq[2] = 0.0f;
            b.dq[3] = 1.0f;
            updateInverseMass(b);
            bodies[i] = b;
        }
        storeMotion(motion, i, Motion{ ld3(b.v), ld3(b.w), ld3(b.dp), ld4(b.dq) });
        storeMass(mass, i, b);
        moving[i] = b.alive && movingBody(b) ? 1 : 0;
    }
}

// src/world.cpp:1604-1621: flags for the active, fast and fast-solid lists.
__global__ void k_classify(const tfs_contact* contacts, uint32_t nc, const tfs_body* bodies, float dt, float speculative,
                           uint32_t* active, uint32_t* fast, uint32_t* fastSolid, uint8_t* colors) {
    for (uint32_t i = gid(); i < nc; i += gstride()) {
        colors[i] = 0xff;
        const tfs_contact& c = contacts[i];
        const tfs_body& a = bodies[c.body_a];
        const tfs_body& b = bodies[c.body_b];
        uint32_t on = 0, f = 0, fs = 0;
        if (movingBody(a) || movingBody(b)) {
            on = 1;
            if (relativeTravel(a, b, dt) > speculative) {
                f = 1;
                fs = !isPlanarLocks(a.locks) && !isPlanarLocks(b.locks) ? 1u : 0u;
            }
        }
        active[i] = on;
        fast[i] = f;
        fastSolid[i] = fs;
    }
}

// prepareContacts (src/solver.cpp:308-313) and the marking half of impactContacts
// (src/world.cpp:1029-1041).
__global__ void __launch_bounds__(kThreads) k_prepare(const uint32_t* active, const DevCounts* cnt, tfs_contact* contacts, const tfs_body* bodies,
                          const float4* motion, const float4* mass, float threshold, uint8_t* struck,
                          uint32_t* points, uint32_t* anyRestitution) {
    const uint32_t n = cnt->nActive;
    uint32_t sum = 0, bounce = 0;
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const uint32_t ci = active[k];
        tfs_contact c = contacts[ci];
        const tfs_body& ta = bodies[c.body_a];
        const tfs_body& tb = bodies[c.body_b];
        prepareContact(viewOf(ta), viewOf(tb), ta, tb, loadMotion(motion, c.body_a), loadMotion(motion, c.body_b),
                       loadMass(mass, c.body_a), loadMass(mass, c.body_b), c);
        contacts[ci] = c;
        sum += (uint32_t)c.count;
        if (c.restitution != 0.0f && c.count > 0) { bounce = 1u; }
        bool hit = false;
        TF_POINTS(i, c.count) { hit = hit || !(c.points[i].relative_velocity >= -threshold); }
        if (hit) {
            if (ta.type == TFS_DYNAMIC) { struck[c.body_a] = 1; }
            if (tb.type == TFS_DYNAMIC) { struck[c.body_b] = 1; }
        }
    }
    waveAdd(points, sum);
    if (__ballot(bounce != 0u) != 0ull && (threadIdx.x % 64) == 0) { anyRestitution[0] = 1u; }
}

// impactContacts' list (src/world.cpp:1042-1048): active contacts that touch a struck body.
__global__ void k_impact_flag(const uint32_t* active, const DevCounts* cnt, const tfs_contact* contacts,
                              const uint8_t* struck, uint32_t* flags) {
    const uint32_t n = cnt->nActive;
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const tfs_contact& c = contacts[active[k]];
        flags[k] = (struck[c.body_a] || struck[c.body_b]) ? 1u : 0u;
    }
}

// Maps positions in the active list to contact indices.
__global__ void k_gather_list(const uint32_t* positions, const uint32_t* nPtr, const uint32_t* active, uint32_t* out) {
    const uint32_t n = *nPtr;
    for (uint32_t k = gid(); k < n; k += gstride()) { out[k] = active[positions[k]]; }
}

__global__ void k_points_of(const uint32_t* list, const uint32_t* nPtr, const tfs_contact* contacts, uint32_t* points) {
    const uint32_t n = *nPtr;
    uint32_t sum = 0;
    for (uint32_t k = gid(); k < n; k += gstride()) { sum += (uint32_t)contacts[list[k]].count; }
    waveAdd(points, sum);
}

// ---- colouring ----------------------------------------------------------------------------------
//
// colorContacts (src/lanes.cpp:356-374) gives each contact, in active order, the lowest colour
// free on both of its moving bodies. A contact's colour depends only on the earlier contacts that
// share a moving body with it, so the kernel runs every contact as soon as its bodies have seen all
// their earlier contacts: each moving body keeps one 64-bit word, (contacts seen << 32) | colours
// taken, and a contact waits until both its words count exactly its rank on that body. The result
// is the sequential greedy colouring, bit for bit.

__device__ __forceinline__ bool colourable(const tfs_contact& c) { return !c.planar && !(c.rolling_resistance > 0.0f); }

__global__ void k_ep_count(const uint32_t* active, const DevCounts* cnt, const tfs_contact* contacts,
                           const uint8_t* moving, uint32_t* deg) {
    const uint32_t n = cnt->nActive;
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const tfs_contact& c = contacts[active[k]];
        if (!colourable(c)) { continue; }
        if (moving[c.body_a]) { atomicAdd(&deg[c.body_a], 1u); }
        if (moving[c.body_b]) { atomicAdd(&deg[c.body_b], 1u); }
    }
}

__global__ void k_ep_fill(const uint32_t* active, const DevCounts* cnt, const tfs_contact* contacts,
                          const uint8_t* moving, const uint32_t* start, uint32_t* cursor, uint32_t* ep) {
    const uint32_t n = cnt->nActive;
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const tfs_contact& c = contacts[active[k]];
        if (!colourable(c)) { continue; }
        if (moving[c.body_a]) { ep[start[c.body_a] + atomicAdd(&cursor[c.body_a], 1u)] = k << 1; }
        if (moving[c.body_b]) { ep[start[c.body_b] + atomicAdd(&cursor[c.body_b], 1u)] = (k << 1) | 1u; }
    }
}

// Sorts each body's endpoints by active position and writes each endpoint's rank on its body.
__global__ void k_ep_rank(uint32_t nb, const uint32_t* deg, const uint32_t* start, uint32_t* ep, uint32_t* rank) {
    for (uint32_t x = gid(); x < nb; x += gstride()) {
        const uint32_t d = deg[x];
        uint32_t* e = ep + start[x];
        for (uint32_t i = 1; i < d; ++i) {
            const uint32_t v = e[i];
            uint32_t j = i;
            while (j > 0 && e[j - 1] > v) {
                e[j] = e[j - 1];
                --j;
            }
            e[j] = v;
        }
        for (uint32_t i = 0; i < d; ++i) { rank[e[i]] = i; }
    }
}

__device__ __forceinline__ unsigned long long loadWord(const unsigned long long* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void storeWord(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Blocks take chunks of the active list in order from a ticket, so every chunk a contact can wait
// on belongs to a block that is already running; the first unfinished contact can always proceed.
//
// A contact waits on the contact before it on each of its moving bodies (the body's endpoint of rank
// r - 1, from the position-sorted endpoints). When that contact is in the same chunk, the wait is in
// LDS: each finished contact leaves its bodies' words in its slot and then marks the slot done, at
// work-group scope, a round trip of about a hundred cycles where a word in global memory, read at agent
// scope across the XCDs, costs a microsecond or more. Most links of the pile's dependency chain are
// within a chunk (contacts come in key order, a body's contacts as the first body together).
constexpr int kColorThreads = 1024;

__global__ void __launch_bounds__(kColorThreads)
    k_color(const uint32_t* active, const DevCounts* cntIn, tfs_contact* contacts, const uint8_t* moving,
            const uint32_t* rank, const uint32_t* epStart, const uint32_t* ep, unsigned long long* word,
            uint8_t* colors, uint32_t* ticket, DevCounts* cnt) {
    __shared__ uint32_t chunk;
    __shared__ uint32_t slotDone[kColorThreads];
    __shared__ unsigned long long slotWord[kColorThreads][2];
    const uint32_t n = cntIn->nActive;
    for (;;) {
        if (threadIdx.x == 0) { chunk = atomicAdd(ticket, 1u); }
        slotDone[threadIdx.x] = 0u;
        __syncthreads();
        const uint32_t base = chunk * blockDim.x;
        const uint32_t k = base + threadIdx.x;
        if (base >= n) { break; }
        if (k < n) {
            const uint32_t ci = active[k];
            const uint32_t a = contacts[ci].body_a, b = contacts[ci].body_b;
            uint8_t col = TFS_COLORS;
            if (colourable(contacts[ci])) {
                const bool ma = moving[a], mb = moving[b];
                const uint32_t ra = rank[k << 1], rb = rank[(k << 1) | 1u];
                // the predecessor's endpoint on each body, when it is in this chunk: its slot and side
                uint32_t la = 0xffffffffu, lb = 0xffffffffu;
                if (ma && ra > 0) {
                    const uint32_t e = ep[epStart[a] + ra - 1];
                    if ((e >> 1) >= base) { la = (((e >> 1) - base) << 1) | (e & 1u); }
                }
                if (mb && rb > 0) {
                    const uint32_t e = ep[epStart[b] + rb - 1];
                    if ((e >> 1) >= base) { lb = (((e >> 1) - base) << 1) | (e & 1u); }
                }
                // First the links from earlier chunks, in global memory: a wave polls them all before
                // it looks at its chunk's own links, so that those then go at LDS speed (a wave's loop
                // turn lasts as long as its slowest lane's load).
                unsigned long long wa = 0ull, wb = 0ull;
                bool readyA = !ma || la != 0xffffffffu, readyB = !mb || lb != 0xffffffffu;
                uint32_t spins = 0;
                while (!(readyA && readyB)) {
                    if (!readyA) {
                        wa = loadWord(&word[a]);
                        readyA = (uint32_t)(wa >> 32) == ra;
                    }
                    if (!readyB) {
                        wb = loadWord(&word[b]);
                        readyB = (uint32_t)(wb >> 32) == rb;
                    }
                    if (readyA && readyB) { break; }
                    if (++spins > kSpinLimit) {
                        atomicOr(&cnt->error, 1u);
                        readyA = readyB = true;
                        la = lb = 0xffffffffu;
                    } else {
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                // Then the links within the chunk, through the slots. The work happens inside the
                // loop: a lane that left the loop would wait for the rest of its wave, which may be
                // waiting for it.
                bool done = false;
                while (!done) {
                    bool okA = true, okB = true;
                    if (ma && la != 0xffffffffu) {
                        okA = __hip_atomic_load(&slotDone[la >> 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u;
                        if (okA) { wa = slotWord[la >> 1][la & 1u]; }
                    }
                    if (mb && lb != 0xffffffffu) {
                        okB = __hip_atomic_load(&slotDone[lb >> 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u;
                        if (okB) { wb = slotWord[lb >> 1][lb & 1u]; }
                    }
                    if (okA && okB) {
                        const uint32_t taken = (uint32_t)wa | (uint32_t)wb;
                        uint32_t add = 0;
                        if (taken != 0xffffffffu) {
                            const uint32_t bit = (uint32_t)__builtin_ctz(~taken);
                            col = (uint8_t)bit;
                            add = 1u << bit;
                        }
                        const unsigned long long na = ((unsigned long long)(ra + 1) << 32) | ((uint32_t)wa | add);
                        const unsigned long long nb = ((unsigned long long)(rb + 1) << 32) | ((uint32_t)wb | add);
                        if (ma) { storeWord(&word[a], na); }
                        if (mb) { storeWord(&word[b], nb); }
                        slotWord[threadIdx.x][0] = na;
                        slotWord[threadIdx.x][1] = nb;
                        __hip_atomic_store(&slotDone[threadIdx.x], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                        done = true;
                    } else if (++spins > kSpinLimit) {
                        atomicOr(&cnt->error, 1u);
                        __hip_atomic_store(&slotDone[threadIdx.x], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                        done = true;
                    } else {
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
            }
            colors[ci] = col;
            contacts[ci].laned = col < TFS_COLORS ? 1 : 0;
        }
        __syncthreads();
    }
}

// ---- row order: batchContacts (src/lanes.cpp:376-405) -------------------------------------------
//
// Colour by colour; within a colour, contacts with more points first; ties and the uncoloured rest
// in list order. A stable counting sort over 5 x 32 + 1 keys.

__device__ __forceinline__ uint32_t rowKey(const tfs_contact& c, uint8_t col) {
    return col >= TFS_COLORS ? 5u * TFS_COLORS : 5u * col + (4u - (uint32_t)c.count);
}

__global__ void k_row_hist(const uint32_t* list, const uint32_t* nPtr, const tfs_contact* contacts, const uint8_t* colors,
                           uint32_t* hist, uint32_t nBlocks) {
    __shared__ uint32_t h[kKeys];
    const uint32_t n = *nPtr;
    for (uint32_t k = threadIdx.x; k < kKeys; k += blockDim.x) { h[k] = 0; }
    __syncthreads();
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        const uint32_t ci = list[k];
        atomicAdd(&h[rowKey(contacts[ci], colors[ci])], 1u);
    }
    __syncthreads();
    for (uint32_t key = threadIdx.x; key < kKeys; key += blockDim.x) { hist[key * nBlocks + blockIdx.x] = h[key]; }
}

__global__ void __launch_bounds__(kThreads) k_row_scatter(const uint32_t* list, const uint32_t* nPtr,
                                                          const tfs_contact* contacts, const uint8_t* colors,
                                                          const uint32_t* offsets, uint32_t nBlocks, uint32_t* out) {
    // Rank of each item among the earlier items of the block with the same key: within the wave from
    // ballots on the key's bits, across the block's waves from per-wave key counts.
    __shared__ uint32_t waveCount[kThreads / 64][kKeys];
    const uint32_t n = *nPtr;
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (uint32_t x = threadIdx.x; x < (kThreads / 64) * kKeys; x += blockDim.x) { (&waveCount[0][0])[x] = 0; }
    __syncthreads();
    const bool on = k < n;
    uint32_t key = kKeys, ci = 0;
    if (on) {
        ci = list[k];
        key = rowKey(contacts[ci], colors[ci]);
    }
    unsigned long long same = __ballot(on);
    for (int bit = 0; bit < 8; ++bit) {
        const unsigned long long b = __ballot((key >> bit) & 1u);
        same &= ((key >> bit) & 1u) ? b : ~b;
    }
    const unsigned long long below = lane ? (~0ull >> (64 - lane)) : 0ull;
    const uint32_t inWave = (uint32_t)__popcll(same & below);
    if (on && inWave == 0) { waveCount[wave][key] = (uint32_t)__popcll(same); }
    __syncthreads();
    if (on) {
        uint32_t before = inWave;
        for (uint32_t w = 0; w < wave; ++w) { before += waveCount[w][key]; }
        out[offsets[key * nBlocks + blockIdx.x] + before] = ci;
    }
}

// Start of each colour's rows (entry 32 is the start of the rest) and each colour's point count: a
// colour's five keys hold its rows with 4, 3, 2, 1 and 0 points.
__global__ void k_row_starts(const uint32_t* offsets, uint32_t nBlocks, uint32_t* starts, uint32_t* points) {
    const uint32_t c = threadIdx.x;
    if (c <= TFS_COLORS) { starts[c] = offsets[(5u * c) * nBlocks]; }
    if (c < TFS_COLORS) {
        uint32_t sum = 0;
        for (uint32_t j = 0; j < 4; ++j) {
            const uint32_t rowsWith = offsets[(5u * c + j + 1) * nBlocks] - offsets[(5u * c + j) * nBlocks];
            sum += (4u - j) * rowsWith;
        }
        points[c] = sum;
    }
}

// ---- rows --------------------------------------------------------------------------------------

// Structure-of-arrays rows: entry [r] or [k * cap + r] for point k. Impact rows are a list of
// active row indices (`at`): the same contacts, solved again in their own order.
struct Rows {
    uint4* hdr;     // body a, body b, flags, contact index
    float4* nrm;    // normal, friction
    float4* tg1;    // tangent1, restitution
    float4* tg2;    // tangent2, 0
    float4* pa;     // anchor a, adjusted separation
    float4* pb;     // anchor b, relative velocity
    float4* imp;    // normal impulse, tangent impulse
    float* peak;    // largest normal impulse
    float4* ax;     // point axes: [((k * 3 + d) * 5 + q) * cap + r], d = normal, tangent1, tangent2
    float4* sep;    // each point's separation in the current pass (one lane per point), written by the pass
    uint32_t cap;
};

// An axis as 5 float4: q0 = linearA, angularA.x; q1 = angularA.yz, linearB.xy; q2 = linearB.z,
// angularB; q3 = armA, armB.x; q4 = armB.yz, mass. A body's side of the axis (what a push applies to
// it) is q0-q1 for body A and q1-q2 for body B, so the warm start reads two float4 per axis.
__device__ __forceinline__ Axis axisOf(float4 a, float4 b, float4 e, float4 f, float4 g) {
    Axis x;
    x.linearA = vec(a.x, a.y, a.z);
    x.angularA = vec(a.w, b.x, b.y);
    x.linearB = vec(b.z, b.w, e.x);
    x.angularB = vec(e.y, e.z, e.w);
    x.armA = vec(f.x, f.y, f.z);
    x.armB = vec(f.w, g.x, g.y);
    x.mass = g.z;
    return x;
}
__device__ __forceinline__ void storeAxis(const Rows& rows, uint32_t r, int k, int d, const Axis& x) {
    float4* q = rows.ax + (size_t)((k * 3 + d) * 5) * rows.cap + r;
    const size_t c = rows.cap;
    q[0] = make_float4(x.linearA.x, x.linearA.y, x.linearA.z, x.angularA.x);
    q[c] = make_float4(x.angularA.y, x.angularA.z, x.linearB.x, x.linearB.y);
    q[2 * c] = make_float4(x.linearB.z, x.angularB.x, x.angularB.y, x.angularB.z);
    q[3 * c] = make_float4(x.armA.x, x.armA.y, x.armA.z, x.armB.x);
    q[4 * c] = make_float4(x.armB.y, x.armB.z, x.mass, 0.0f);
}
__device__ __forceinline__ Axis loadAxis(const Rows& rows, uint32_t r, int k, int d) {
    const float4* q = rows.ax + (size_t)((k * 3 + d) * 5) * rows.cap + r;
    const size_t c = rows.cap;
    return axisOf(q[0], q[c], q[2 * c], q[3 * c], q[4 * c]);
}

struct LoadedAxes {
    const Rows& rows;
    uint32_t r;
    __device__ __forceinline__ Axis operator()(int i, int d) const { return loadAxis(rows, r, i, d); }
};

// Packs the coloured contacts into rows in row order and builds their point axes from the bodies'
// current masses, as prepareLanes does (src/lanes.cpp:66-110, 409-411).
// The contact's points are read before anything is written, so one row waits for memory once.
// Row r from contact ci: header, directions, points and their three axes.
__device__ __forceinline__ void packRow(uint32_t r, uint32_t ci, const tfs_contact* __restrict__ contacts,
                                        const uint8_t* __restrict__ moving, const float4* __restrict__ mass,
                                        const Rows& rows, uint32_t* __restrict__ rowOf) {
    {
        const tfs_contact& c = contacts[ci];
        const int count = c.count;
        tfs_point pt[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < count) { pt[k] = c.points[k]; }
        }
        const uint32_t flags = (uint32_t)count | (moving[c.body_a] ? kMoveA : 0u) | (moving[c.body_b] ? kMoveB : 0u);
        const Mass ma = loadMass(mass, c.body_a), mb = loadMass(mass, c.body_b);
        const V3 dir[3] = { ld3(c.normal), ld3(c.tangent1), ld3(c.tangent2) };
        rows.hdr[r] = make_uint4(c.body_a, c.body_b, flags, ci);
        rows.nrm[r] = make_float4(c.normal[0], c.normal[1], c.normal[2], c.friction);
        rows.tg1[r] = make_float4(c.tangent1[0], c.tangent1[1], c.tangent1[2], c.restitution);
        rows.tg2[r] = make_float4(c.tangent2[0], c.tangent2[1], c.tangent2[2], 0.0f);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k >= count) { break; }
            const tfs_point& p = pt[k];
            const uint32_t j = (uint32_t)k * rows.cap + r;
            rows.pa[j] = make_float4(p.anchor_a[0], p.anchor_a[1], p.anchor_a[2], p.adjusted_separation);
            rows.pb[j] = make_float4(p.anchor_b[0], p.anchor_b[1], p.anchor_b[2], p.relative_velocity);
            rows.imp[j] = make_float4(p.normal_impulse, p.tangent_impulse[0], p.tangent_impulse[1], p.tangent_impulse[2]);
            rows.peak[j] = p.max_normal_impulse;
#pragma unroll
            for (int d = 0; d < 3; ++d) { storeAxis(rows, r, k, d, makeAxis(ma, mb, ld3(p.anchor_a), ld3(p.anchor_b), dir[d])); }
        }
        rowOf[ci] = r;
    }
}

__global__ void __launch_bounds__(kThreads) k_pack(const uint32_t* __restrict__ list, uint32_t n, const tfs_contact* __restrict__ contacts,
                       const uint8_t* __restrict__ moving, const float4* __restrict__ mass, Rows rows,
                       uint32_t* __restrict__ rowOf) {
    for (uint32_t r = gid(); r < n; r += gstride()) { packRow(r, list[r], contacts, moving, mass, rows, rowOf); }
}

// The impact rows' headers: their active row's header with the active row index in place of the
// contact index.
__global__ void k_impact_rows(const uint32_t* list, uint32_t n, const uint32_t* rowOf, const uint4* hdr, uint4* impactHdr) {
    for (uint32_t i = gid(); i < n; i += gstride()) {
        const uint32_t r = rowOf[list[i]];
        const uint4 h = hdr[r];
        impactHdr[i] = make_uint4(h.x, h.y, h.z, r);
    }
}

// Impulses of the coloured rows back into the contacts.
__global__ void k_unpack(uint32_t n, Rows rows, tfs_contact* contacts) {
    for (uint32_t r = gid(); r < n; r += gstride()) {
        const uint4 h = rows.hdr[r];
        tfs_contact& c = contacts[h.w];
        const int count = (int)(h.z & kCountMask);
        for (int k = 0; k < count; ++k) {
            const uint32_t j = (uint32_t)k * rows.cap + r;
            const float4 im = rows.imp[j];
            c.points[k].normal_impulse = im.x;
            c.points[k].tangent_impulse[0] = im.y;
            c.points[k].tangent_impulse[1] = im.z;
            c.points[k].tangent_impulse[2] = im.w;
            c.points[k].max_normal_impulse = rows.peak[j];
        }
    }
}

// Velocity and impulse accesses of the rows. Within one launch per colour, kernel boundaries make
// every write visible to the next colour, and plain loads and stores serve. In a grid-barrier pass
// the colours follow each other inside one launch, across blocks on different XCDs; there the rows
// read and write velocities and impulses with device-coherent (agent-scope) atomic loads and stores,
// so that a barrier only has to order them, not flush caches.
template <bool COHERENT>
struct Access {
    __device__ __forceinline__ static float4 ld(const float4* p) {
        if (!COHERENT) { return *p; }
        const unsigned long long* q = (const unsigned long long*)p;
        const unsigned long long a = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long b = __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        float4 r;
        memcpy(&r.x, &a, 8);
        memcpy(&r.z, &b, 8);
        return r;
    }
    __device__ __forceinline__ static void st(float4* p, float4 v) {
        if (!COHERENT) {
            *p = v;
            return;
        }
        unsigned long long a, b;
        memcpy(&a, &v.x, 8);
        memcpy(&b, &v.z, 8);
        unsigned long long* q = (unsigned long long*)p;
        __hip_atomic_store(q, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(q + 1, b, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __device__ __forceinline__ static float ld1(const float* p) {
        if (!COHERENT) { return *p; }
        const uint32_t a = __hip_atomic_load((const uint32_t*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        float r;
        memcpy(&r, &a, 4);
        return r;
    }
    __device__ __forceinline__ static void st1(float* p, float v) {
        if (!COHERENT) {
            *p = v;
            return;
        }
        uint32_t a;
        memcpy(&a, &v, 4);
        __hip_atomic_store((uint32_t*)p, a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// The packed axes of a row, already in registers: axv[point][axis * 5 + q].
template <int Q>
struct RegAxes {
    const float4 (&axv)[4][Q];
    __device__ __forceinline__ Axis operator()(int i, int d) const {
        return axisOf(axv[i][d * 5], axv[i][d * 5 + 1], axv[i][d * 5 + 2], axv[i][d * 5 + 3], axv[i][d * 5 + 4]);
    }
};

// The axes of a row, built on the spot from the bodies' masses (the same bits as the packed ones).
struct RowMadeAxes {
    Mass ma, mb;
    const Row& r;
    const RowPoint* p;
    __device__ __forceinline__ Axis operator()(int i, int d) const {
        return makeAxis(ma, mb, p[i].ra, p[i].rb, d == 0 ? r.n : d == 1 ? r.t1 : r.t2);
    }
};

// What a recording pass keeps of each row it solves, for the reverse sweep of the fixed-contact
// model: the velocities the row saw (v[q * cap + r], q = vA, wA, vB, wB), its points' impulses
// (imp[k * cap + r]) and separations (sep[r], one lane per point). Rows are indexed as in Rows.
struct Tape {
    float4* v;
    float4* imp;
    float4* sep;
    uint32_t cap;
};

// One row, r, with header h (bodies, flags). Rows of one colour move no body in common
// (src/internal.h:607-609), so within a colour each row's thread owns the row's two moving bodies.
// MADE builds the point axes on the spot from the masses instead of reading the packed ones.
template <int MODE, bool COHERENT, bool MADE, bool TAPE>
__device__ __forceinline__ void rowStep(const Rows& rows, uint32_t r, uint4 h, float4* motion, const float4* mass,
                                        const Ctx& ctx, const Tape& tape) {
    using X = Access<COHERENT>;
    const int count = (int)(h.z & kCountMask);
    if (count == 0) { return; }
    const float4 n = rows.nrm[r], t1 = rows.tg1[r], t2 = rows.tg2[r];
    const Row row{ h.z, vec(n.x, n.y, n.z), vec(t1.x, t1.y, t1.z), vec(t2.x, t2.y, t2.z), n.w, t1.w };
    if (MODE == kRestitute && row.restitution == 0.0f) { return; }
    // Every load the row needs is issued here, before any arithmetic, so the row waits for memory
    // once rather than once per axis.
    RowPoint p[4];
    float4 axv[4][MADE ? 1 : 15];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < count) {
            const uint32_t j = (uint32_t)k * rows.cap + r;
            const float4 a = rows.pa[j], b = rows.pb[j], im = X::ld(&rows.imp[j]);
            p[k].ra = vec(a.x, a.y, a.z);
            p[k].adj = a.w;
            p[k].rb = vec(b.x, b.y, b.z);
            p[k].rv = b.w;
            p[k].ni = im.x;
            p[k].ti = vec(im.y, im.z, im.w);
            p[k].nimax = X::ld1(&rows.peak[j]);
            if (!MADE) {
#pragma unroll
                for (int q = 0; q < 15; ++q) { axv[k][q] = rows.ax[(size_t)(k * 15 + q) * rows.cap + r]; }
            }
        }
    }
    const float4 va = X::ld(&motion[4 * h.x]), wa = X::ld(&motion[4 * h.x + 1]);
    const float4 vb = X::ld(&motion[4 * h.y]), wb = X::ld(&motion[4 * h.y + 1]);
    const float4 pa = motion[4 * h.x + 2], qa = motion[4 * h.x + 3], pb = motion[4 * h.y + 2], qb = motion[4 * h.y + 3];
    Motion A{ vec(va.x, va.y, va.z), vec(wa.x, wa.y, wa.z), vec(pa.x, pa.y, pa.z), quat(qa.x, qa.y, qa.z, qa.w) };
    Motion B{ vec(vb.x, vb.y, vb.z), vec(wb.x, wb.y, wb.z), vec(pb.x, pb.y, pb.z), quat(qb.x, qb.y, qb.z, qb.w) };
    if (TAPE && tape.v) {
        const uint32_t c = tape.cap;
        tape.v[r] = va;
        tape.v[c + r] = wa;
        tape.v[2 * c + r] = vb;
        tape.v[3 * c + r] = wb;
        const V3 dp = B.dp - A.dp;
        float sp[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < count) {
                tape.imp[(uint32_t)k * c + r] = make_float4(p[k].ni, p[k].ti.x, p[k].ti.y, p[k].ti.z);
                sp[k] = pointSeparation(A, B, dp, p[k], row.n);
            }
        }
        tape.sep[r] = make_float4(sp[0], sp[1], sp[2], sp[3]);
    }
    if (MADE) {
        const RowMadeAxes ax{ loadMass(mass, h.x), loadMass(mass, h.y), row, p };
        tf::rows<MODE>(row, p, ax, A, B, ctx);
    } else {
        const RegAxes<MADE ? 1 : 15> ax{ axv };
        tf::rows<MODE>(row, p, ax, A, B, ctx);
    }
    if (MODE != kWarm) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < count) {
                const uint32_t j = (uint32_t)k * rows.cap + r;
                X::st(&rows.imp[j], make_float4(p[k].ni, p[k].ti.x, p[k].ti.y, p[k].ti.z));
                X::st1(&rows.peak[j], p[k].nimax);
            }
        }
    }
    if (h.z & kMoveA) {
        X::st(&motion[4 * h.x], make_float4(A.v.x, A.v.y, A.v.z, 0.0f));
        X::st(&motion[4 * h.x + 1], make_float4(A.w.x, A.w.y, A.w.z, 0.0f));
    }
    if (h.z & kMoveB) {
        X::st(&motion[4 * h.y], make_float4(B.v.x, B.v.y, B.v.z, 0.0f));
        X::st(&motion[4 * h.y + 1], make_float4(B.w.x, B.w.y, B.w.z, 0.0f));
    }
}

// One colour per launch; impactHdr is set for impact rows.
template <int MODE, bool MADE>
__global__ void __launch_bounds__(kThreads) k_rows(Rows rows, const uint4* impactHdr, uint32_t begin, uint32_t end,
                                                   float4* motion, const float4* mass, Ctx ctx, Tape tape) {
    const uint32_t i = begin + gid();
    if (i >= end) { return; }
    const uint4 h = impactHdr ? impactHdr[i] : rows.hdr[i];
    rowStep<MODE, false, MADE, true>(rows, impactHdr ? h.w : i, h, motion, mass, ctx, impactHdr ? Tape{} : tape);
}

// A whole pass in one launch with a grid barrier: the colours of the active rows, then those of the impact
// rows, one after the other with a grid-wide barrier between them.
constexpr int kPassThreads = 256;
constexpr int kMaxSegments = 2 * TFS_COLORS;

struct PassPlan {
    uint32_t n;                 // segments
    uint32_t impactFrom;        // segments from here are impact rows
    uint32_t tailFrom;          // segments from here are solved by block 0 alone (passTail)
    uint32_t sepRows;           // rows 0..sepRows-1 get their separations in the pass's first phase
    uint32_t begin[kMaxSegments], end[kMaxSegments];
};

// The grid-wide barrier of a pass launch (launchGrid). Blocks arrive in kBarrierGroups groups (block b in
// group b % kBarrierGroups), each group on a counter of its own cache line, with an atomic add whose
// result nobody waits for; then the first wave of every block reads all the group counters at once,
// one lane per group, until every group is complete. Counters only grow: a launch is given each
// counter's value before it (BarrierBase) and knows how much each barrier adds.
constexpr uint32_t kBarrierGroups = 8;
struct alignas(128) BarrierLine {
    unsigned long long n;
    unsigned long long pad_[15];
};
struct BarrierLines {
    BarrierLine group[kBarrierGroups];
};
struct BarrierBase {
    unsigned long long group[kBarrierGroups];
};

__host__ __device__ __forceinline__ uint32_t groupBlocks(uint32_t grid, uint32_t gi) {
    return gi < grid ? (grid - gi + kBarrierGroups - 1) / kBarrierGroups : 0u;
}

// What lane l of a block's first wave waits for: group l's count after the next barrier (lanes past
// the last group watch nothing).
struct BarrierPhase {
    unsigned long long target;
    uint32_t step;
    __device__ __forceinline__ explicit BarrierPhase(const BarrierBase& b) {
        const uint32_t lane = threadIdx.x & 63u;
        step = lane < kBarrierGroups ? groupBlocks(gridDim.x, lane) : 0u;
        target = step ? b.group[lane] : 0ull;
    }
};

// The base of the launch after one with `barriers` barriers over `grid` blocks.
inline void advanceBarrier(BarrierBase& b, uint32_t grid, uint32_t barriers) {
    for (uint32_t gi = 0; gi < kBarrierGroups; ++gi) { b.group[gi] += (unsigned long long)barriers * groupBlocks(grid, gi); }
}

// Every block waits for its own coherent stores to complete, then arrives (gridArrive) and waits
// until every block has (gridWait). Rows exchange data only through coherent accesses, so no cache
// needs flushing.
__device__ __forceinline__ void gridArrive(BarrierLines* bar, BarrierPhase& phase) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) { atomicAdd(&bar->group[blockIdx.x % kBarrierGroups].n, 1ull); }
    phase.target += phase.step;
}

__device__ __forceinline__ void gridWait(BarrierLines* bar, const BarrierPhase& phase, DevCounts* cnt) {
    if (threadIdx.x < 64) {
        const uint32_t lane = threadIdx.x;
        uint32_t spins = 0;
        for (;;) {
            const bool done = phase.step == 0 ||
                              __hip_atomic_load(&bar->group[lane].n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= phase.target;
            if (__all(done)) { break; }
            if (++spins > kSpinLimit) {
                if (lane == 0) { atomicOr(&cnt->error, 2u); }
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
    }
    __syncthreads();
}

__device__ __forceinline__ void gridSync(BarrierLines* bar, BarrierPhase& phase, DevCounts* cnt) {
    gridArrive(bar, phase);
    gridWait(bar, phase, cnt);
}

// A thread's slot in a pass over a list: slot t serves list entry begin + t. Consecutive slots form
// waves (64 entries that load together), and consecutive waves go to different blocks, so a short
// list spreads over as many compute units as the grid has blocks.
__device__ __forceinline__ uint32_t passSlot() {
    return ((threadIdx.x >> 6) * gridDim.x + blockIdx.x) * 64u + (threadIdx.x & 63u);
}

// ---- the tail of a pass: its last small colours in one block ----------------------------------------
//
// The last colours of a pass are often a few rows each (the last active colours, and the impact rows'
// colours), and each would cost a grid barrier plus a full round trip to memory for its rows. The
// tail hands them to block 0 alone: every row of the tail gets one thread, which loads the row's data
// into registers once and keeps its impulses there across the tail (an impact row whose active row
// is also in the tail is the same thread's second visit); the bodies' velocities live in LDS slots
// (a hash of body index to slot) shared by the tail's threads, and colours follow each other with
// __syncthreads. The other blocks leave after the last grid barrier without waiting at it. Every
// row's operations and their order are those of the grid path, so the results are the same bits.

constexpr uint32_t kTailSlots = 512;      // LDS body slots (at most 2 per tail row)
constexpr uint32_t kTailSegRows = 64;     // a tail colour has at most one wave of rows
constexpr uint32_t kNoSlot = 0xffffffffu;

// What a row reads that the pass does not change: directions, its points' separations in this pass,
// and the packed axes (MADE: the anchors, to make the axes from).
template <bool MADE>
struct RowData {
    float4 n, t1, t2;
    float4 sep;
    float4 pa[MADE ? 4 : 1], pb[MADE ? 4 : 1];
    float4 axv[4][MADE ? 1 : 15];
};

// Everything but the separations, which the pass's first phase writes (loadRowSep).
template <bool MADE>
__device__ __forceinline__ void loadRowData(const Rows& rows, uint32_t r, uint4 h, RowData<MADE>& d) {
    const int count = (int)(h.z & kCountMask);
    d.n = rows.nrm[r];
    d.t1 = rows.tg1[r];
    d.t2 = rows.tg2[r];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < count) {
            const uint32_t j = (uint32_t)k * rows.cap + r;
            if constexpr (MADE) {
                d.pa[k] = rows.pa[j];
                d.pb[k] = rows.pb[j];
            } else {
#pragma unroll
                for (int q = 0; q < 15; ++q) { d.axv[k][q] = rows.ax[(size_t)(k * 15 + q) * rows.cap + r]; }
            }
        }
    }
}

// A row's separations, written in the same launch by the pass's first phase on any block: loaded
// coherently, once that phase's barrier has passed.
__device__ __forceinline__ float4 loadRowSep(const Rows& rows, uint32_t r) { return Access<true>::ld(&rows.sep[r]); }

// The first phase of a pass: every point's current separation (solveLanes, src/lanes.cpp:236-240: its
// anchors turned by the bodies' rotation this step, plus their travel dp = b.dp - a.dp, along the
// normal, plus the adjusted separation), the same operations as pointSeparation in the row solve. It
// depends only on the bodies' dp and dq, which no row of a pass changes, so the rows of every colour
// take it from here instead of computing it in their chain. Stored coherently for the threads that
// solve the row, on any block. DPQ: dp and dq were written in the same launch (the fused substep
// kernel), by other blocks: they are loaded coherently.
template <bool DPQ>
__device__ __forceinline__ void rowSeparations(const Rows& rows, uint32_t r, const float4* motion) {
    using X = Access<DPQ>;
    const uint4 h = rows.hdr[r];
    const int count = (int)(h.z & kCountMask);
    const float4 n4 = rows.nrm[r];
    const float4 dpA = X::ld(&motion[4 * h.x + 2]), dqA = X::ld(&motion[4 * h.x + 3]);
    const float4 dpB = X::ld(&motion[4 * h.y + 2]), dqB = X::ld(&motion[4 * h.y + 3]);
    const Motion A{ vec(0.0f, 0.0f, 0.0f), vec(0.0f, 0.0f, 0.0f), vec(dpA.x, dpA.y, dpA.z), quat(dqA.x, dqA.y, dqA.z, dqA.w) };
    const Motion B{ vec(0.0f, 0.0f, 0.0f), vec(0.0f, 0.0f, 0.0f), vec(dpB.x, dpB.y, dpB.z), quat(dqB.x, dqB.y, dqB.z, dqB.w) };
    const V3 dp = B.dp - A.dp, n = vec(n4.x, n4.y, n4.z);
    float sp[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < count) {
            const uint32_t j = (uint32_t)k * rows.cap + r;
            const float4 a = rows.pa[j], b = rows.pb[j];
            RowPoint p;
            p.ra = vec(a.x, a.y, a.z);
            p.adj = a.w;
            p.rb = vec(b.x, b.y, b.z);
            sp[k] = pointSeparation(A, B, dp, p, n);
        }
    }
    Access<true>::st(&rows.sep[r], make_float4(sp[0], sp[1], sp[2], sp[3]));
}

// rowStep over data already loaded (loadRowData): only what rows of the pass write (velocities,
// impulses) is loaded here, coherently. The same operations on the same values as rowStep.
template <int MODE, bool MADE, bool TAPE>
__device__ __forceinline__ void rowSolveLoaded(const Rows& rows, uint32_t r, uint4 h, const RowData<MADE>& d, float4* motion,
                                               const float4* mass, const Ctx& ctx, const Tape& tape) {
    using X = Access<true>;
    const int count = (int)(h.z & kCountMask);
    if (count == 0) { return; }
    const Row row{ h.z, vec(d.n.x, d.n.y, d.n.z), vec(d.t1.x, d.t1.y, d.t1.z), vec(d.t2.x, d.t2.y, d.t2.z), d.n.w, d.t1.w };
    const float sp[4] = { d.sep.x, d.sep.y, d.sep.z, d.sep.w };
    RowPoint p[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < count) {
            const uint32_t j = (uint32_t)k * rows.cap + r;
            const float4 im = X::ld(&rows.imp[j]);
            if constexpr (MADE) {
                p[k].ra = vec(d.pa[k].x, d.pa[k].y, d.pa[k].z);
                p[k].rb = vec(d.pb[k].x, d.pb[k].y, d.pb[k].z);
            }
            p[k].sep = sp[k];
            p[k].ni = im.x;
            p[k].ti = vec(im.y, im.z, im.w);
            p[k].nimax = X::ld1(&rows.peak[j]);
        }
    }
    const float4 va = X::ld(&motion[4 * h.x]), wa = X::ld(&motion[4 * h.x + 1]);
    const float4 vb = X::ld(&motion[4 * h.y]), wb = X::ld(&motion[4 * h.y + 1]);
    // (dp and dq enter only through the separations)
    Motion A{ vec(va.x, va.y, va.z), vec(wa.x, wa.y, wa.z), vec(0.0f, 0.0f, 0.0f), quat(0.0f, 0.0f, 0.0f, 1.0f) };
    Motion B{ vec(vb.x, vb.y, vb.z), vec(wb.x, wb.y, wb.z), vec(0.0f, 0.0f, 0.0f), quat(0.0f, 0.0f, 0.0f, 1.0f) };
    if (TAPE && tape.v) {
        const uint32_t c = tape.cap;
        tape.v[r] = va;
        tape.v[c + r] = wa;
        tape.v[2 * c + r] = vb;
        tape.v[3 * c + r] = wb;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < count) { tape.imp[(uint32_t)k * c + r] = make_float4(p[k].ni, p[k].ti.x, p[k].ti.y, p[k].ti.z); }
        }
        tape.sep[r] = d.sep;
    }
    if constexpr (MADE) {
        const RowMadeAxes ax{ loadMass(mass, h.x), loadMass(mass, h.y), row, p };
        tf::rows<MODE, RowMadeAxes, true>(row, p, ax, A, B, ctx);
    } else {
        const RegAxes<15> ax{ d.axv };
        tf::rows<MODE, RegAxes<15>, true>(row, p, ax, A, B, ctx);
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < count) {
            const uint32_t j = (uint32_t)k * rows.cap + r;
            X::st(&rows.imp[j], make_float4(p[k].ni, p[k].ti.x, p[k].ti.y, p[k].ti.z));
            X::st1(&rows.peak[j], p[k].nimax);
        }
    }
    if (h.z & kMoveA) {
        X::st(&motion[4 * h.x], make_float4(A.v.x, A.v.y, A.v.z, 0.0f));
        X::st(&motion[4 * h.x + 1], make_float4(A.w.x, A.w.y, A.w.z, 0.0f));
    }
    if (h.z & kMoveB) {
        X::st(&motion[4 * h.y], make_float4(B.v.x, B.v.y, B.v.z, 0.0f));
        X::st(&motion[4 * h.y + 1], make_float4(B.w.x, B.w.y, B.w.z, 0.0f));
    }
}

__device__ __forceinline__ uint32_t tailSlot(uint32_t* key, uint32_t body) {
    uint32_t k = (body * 2654435761u) >> 23;   // 9 bits: kTailSlots = 512
    for (uint32_t probe = 0; probe < kTailSlots; ++probe, k = (k + 1) & (kTailSlots - 1)) {
        const uint32_t old = atomicCAS(&key[k], kNoSlot, body);
        if (old == kNoSlot || old == body) { return k; }
    }
    return kNoSlot;
}

// Block 0's part of a pass from segment plan.tailFrom on. With `wait`, the grid phase before it
// ended with gridArrive: data that no row of the pass writes is loaded before gridWait, velocities and
// impulses after it.
// FUSED: the pass runs inside the fused substep kernel, whose later phases read what the tail writes
// from other blocks, so the tail's results go back with coherent stores.
template <int MODE, bool MADE, bool TAPE, bool FUSED = false>
__device__ __forceinline__ void passTail(const Rows& rows, const uint4* impactHdr, const PassPlan& plan, float4* motion,
                                         const float4* mass, const Ctx& ctx, const Tape& tape, BarrierLines* bar,
                                         const BarrierPhase& phase, DevCounts* cnt, bool wait) {
    using X = Access<true>;
    __shared__ uint32_t slotKey[kTailSlots];
    __shared__ float4 slotV[kTailSlots], slotW[kTailSlots];
    __shared__ uint32_t forward[kPassThreads];
    const uint32_t t = threadIdx.x, tf = plan.tailFrom;
    const bool activeTail = tf < plan.impactFrom;
    const uint32_t atBegin = activeTail ? plan.begin[tf] : 0u, atEnd = activeTail ? plan.end[plan.impactFrom - 1] : 0u;
    const uint32_t firstImpact = tf > plan.impactFrom ? tf : plan.impactFrom;
    const uint32_t itBegin = firstImpact < plan.n ? plan.begin[firstImpact] : 0u;
    const uint32_t itEnd = firstImpact < plan.n ? plan.end[plan.n - 1] : 0u;
    const uint32_t atCount = atEnd - atBegin, itCount = itEnd - itBegin;
    for (uint32_t k = t; k < kTailSlots; k += blockDim.x) { slotKey[k] = kNoSlot; }
    forward[t] = kNoSlot;
    __syncthreads();
    // Ownership: thread t < atCount owns active row atBegin + t; thread atCount + k the k-th impact
    // visit, unless its active row is in the tail, whose owner then also takes the visit.
    uint32_t row = kNoSlot, activeSeg = kNoSlot, impactSeg = kNoSlot;
    uint4 h = make_uint4(0u, 0u, 0u, 0u);
    if (t < atCount) {
        row = atBegin + t;
        h = rows.hdr[row];
        for (uint32_t s = tf; s < plan.impactFrom; ++s) {
            if (row >= plan.begin[s] && row < plan.end[s]) { activeSeg = s; }
        }
    } else if (t < atCount + itCount) {
        const uint32_t i = itBegin + (t - atCount);
        const uint4 hi = impactHdr[i];
        uint32_t seg = kNoSlot;
        for (uint32_t s = firstImpact; s < plan.n; ++s) {
            if (i >= plan.begin[s] && i < plan.end[s]) { seg = s; }
        }
        if (hi.w >= atBegin && hi.w < atEnd) {
            forward[hi.w - atBegin] = seg;
        } else {
            row = hi.w;
            h = hi;
            impactSeg = seg;
        }
    }
    __syncthreads();
    if (t < atCount) { impactSeg = forward[t]; }
    const int count = (int)(h.z & kCountMask);
    RowData<MADE> d;
    uint32_t sa = kNoSlot, sb = kNoSlot;
    if (row != kNoSlot) {
        // (the separations are final: the pass's first phase ended with a barrier that every block
        // has passed before any tail begins)
        loadRowData<MADE>(rows, row, h, d);
        d.sep = loadRowSep(rows, row);
        sa = tailSlot(slotKey, h.x);
        sb = tailSlot(slotKey, h.y);
        if (sa == kNoSlot || sb == kNoSlot) { atomicOr(&cnt->error, 4u); }
    }
    // (the slot table holds every body of the tail, at most two per thread, so it cannot fill up)
    const bool owner = row != kNoSlot && sa != kNoSlot && sb != kNoSlot;
    if (wait) {
        gridWait(bar, phase, cnt);
    } else {
        __syncthreads();
    }
    // Velocities and impulses, written by any block before the barrier: coherent loads.
    for (uint32_t k = t; k < kTailSlots; k += blockDim.x) {
        const uint32_t b = slotKey[k];
        if (b != kNoSlot) {
            slotV[k] = X::ld(&motion[4 * b]);
            slotW[k] = X::ld(&motion[4 * b + 1]);
        }
    }
    float4 im[4];
    float peak[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (owner && k < count) {
            const uint32_t j = (uint32_t)k * rows.cap + row;
            im[k] = X::ld(&rows.imp[j]);
            peak[k] = X::ld1(&rows.peak[j]);
        }
    }
    __syncthreads();
    const Row rowv{ h.z, vec(d.n.x, d.n.y, d.n.z), vec(d.t1.x, d.t1.y, d.t1.z), vec(d.t2.x, d.t2.y, d.t2.z), d.n.w, d.t1.w };
    const float sp[4] = { d.sep.x, d.sep.y, d.sep.z, d.sep.w };
    for (uint32_t s = tf; s < plan.n; ++s) {
        if (owner && count > 0 && (s == activeSeg || s == impactSeg)) {
            RowPoint p[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (k < count) {
                    if constexpr (MADE) {
                        p[k].ra = vec(d.pa[k].x, d.pa[k].y, d.pa[k].z);
                        p[k].rb = vec(d.pb[k].x, d.pb[k].y, d.pb[k].z);
                    }
                    p[k].sep = sp[k];
                    p[k].ni = im[k].x;
                    p[k].ti = vec(im[k].y, im[k].z, im[k].w);
                    p[k].nimax = peak[k];
                }
            }
            const float4 va = slotV[sa], wa = slotW[sa], vb = slotV[sb], wb = slotW[sb];
            Motion A{ vec(va.x, va.y, va.z), vec(wa.x, wa.y, wa.z), vec(0.0f, 0.0f, 0.0f), quat(0.0f, 0.0f, 0.0f, 1.0f) };
            Motion B{ vec(vb.x, vb.y, vb.z), vec(wb.x, wb.y, wb.z), vec(0.0f, 0.0f, 0.0f), quat(0.0f, 0.0f, 0.0f, 1.0f) };
            if (TAPE && tape.v && s == activeSeg) {
                const uint32_t c = tape.cap;
                tape.v[row] = va;
                tape.v[c + row] = wa;
                tape.v[2 * c + row] = vb;
                tape.v[3 * c + row] = wb;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    if (k < count) { tape.imp[(uint32_t)k * c + row] = im[k]; }
                }
                tape.sep[row] = d.sep;
            }
            if constexpr (MADE) {
                const RowMadeAxes ax{ loadMass(mass, h.x), loadMass(mass, h.y), rowv, p };
                tf::rows<MODE, RowMadeAxes, true>(rowv, p, ax, A, B, ctx);
            } else {
                const RegAxes<15> ax{ d.axv };
                tf::rows<MODE, RegAxes<15>, true>(rowv, p, ax, A, B, ctx);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (k < count) {
                    im[k] = make_float4(p[k].ni, p[k].ti.x, p[k].ti.y, p[k].ti.z);
                    peak[k] = p[k].nimax;
                }
            }
            if (h.z & kMoveA) {
                slotV[sa] = make_float4(A.v.x, A.v.y, A.v.z, 0.0f);
                slotW[sa] = make_float4(A.w.x, A.w.y, A.w.z, 0.0f);
            }
            if (h.z & kMoveB) {
                slotV[sb] = make_float4(B.v.x, B.v.y, B.v.z, 0.0f);
                slotW[sb] = make_float4(B.w.x, B.w.y, B.w.z, 0.0f);
            }
        }
        __syncthreads();
    }
    // Back to memory: the impulses of every tail row, the velocities of every slot.
    using Y = Access<FUSED>;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (owner && k < count) {
            const uint32_t j = (uint32_t)k * rows.cap + row;
            Y::st(&rows.imp[j], im[k]);
            Y::st1(&rows.peak[j], peak[k]);
        }
    }
    for (uint32_t k = t; k < kTailSlots; k += blockDim.x) {
        const uint32_t b = slotKey[k];
        if (b != kNoSlot) {
            Y::st(&motion[4 * b], slotV[k]);
            Y::st(&motion[4 * b + 1], slotW[k]);
        }
    }
}

// First, with plan.sepRows, every active row's separations (rowSeparations) over all threads, and a
// grid barrier. Headers, directions, axes and separations do not change during the colours, so each
// thread loads them for its first row of the next segment while it waits at the barrier; after the
// barrier only the velocities and impulses that earlier segments wrote are loaded.
// Segments from plan.tailFrom on are the tail (passTail).
// The pass as a device function: k_pass launches it alone; the fused substep kernel (k_substep_fused)
// runs it between its other phases (FUSED), with the grid barrier's phase carried across them.
template <int MODE, bool MADE, bool TAPE, bool FUSED>
__device__ __forceinline__ void passRun(const Rows& rows, const uint4* impactHdr, const PassPlan& plan, float4* motion,
                                        const float4* mass, const Ctx& ctx, BarrierLines* bar, BarrierPhase& phase,
                                        DevCounts* cnt, const Tape& tape) {
    const uint32_t g = passSlot(), stride = gstride();
    auto header = [&](uint32_t s, uint32_t i) { return s >= plan.impactFrom ? impactHdr[i] : rows.hdr[i]; };
    const uint32_t grid = plan.tailFrom;
    if (plan.sepRows > 0) {
        for (uint32_t r = gid(); r < plan.sepRows; r += stride) { rowSeparations<FUSED>(rows, r, motion); }
        gridArrive(bar, phase);
    }
    uint4 next = make_uint4(0u, 0u, 0u, 0u);
    RowData<MADE> nd;
    const bool firstRow = grid > 0 && plan.begin[0] + g < plan.end[0];
    uint32_t firstAt = 0;   // the row whose data the first segment's row reads (an impact row: its active row)
    if (firstRow) {
        next = header(0, plan.begin[0] + g);
        firstAt = plan.impactFrom == 0 ? next.w : plan.begin[0] + g;
        loadRowData<MADE>(rows, firstAt, next, nd);
    }
    if (plan.sepRows > 0) { gridWait(bar, phase, cnt); }
    if (firstRow) { nd.sep = loadRowSep(rows, firstAt); }
    for (uint32_t s = 0; s < grid; ++s) {
        const bool impact = s >= plan.impactFrom;
        const uint32_t first = plan.begin[s] + g;
        const uint4 firstHdr = next;
        const bool more = s + 1 < grid && plan.begin[s + 1] + g < plan.end[s + 1];
        if (more) { next = header(s + 1, plan.begin[s + 1] + g); }
        for (uint32_t i = first; i < plan.end[s]; i += stride) {
            uint4 h = firstHdr;
            if (i != first) {
                h = header(s, i);
                loadRowData<MADE>(rows, impact ? h.w : i, h, nd);
                nd.sep = loadRowSep(rows, impact ? h.w : i);
            }
            rowSolveLoaded<MODE, MADE, TAPE>(rows, impact ? h.w : i, h, nd, motion, mass, ctx, impact ? Tape{} : tape);
        }
        if (s + 1 < grid) {
            // The next segment's first row: what no row of the pass writes is loaded between arriving
            // at the barrier (which waits for this thread's memory operations) and waiting at it.
            gridArrive(bar, phase);
            if (more) {
                const uint32_t at = s + 1 >= plan.impactFrom ? next.w : plan.begin[s + 1] + g;
                loadRowData<MADE>(rows, at, next, nd);
                nd.sep = loadRowSep(rows, at);
            }
            gridWait(bar, phase, cnt);
        } else if (grid < plan.n) {
            gridArrive(bar, phase);
        }
    }
    if (grid < plan.n && blockIdx.x == 0) {
        passTail<MODE, MADE, TAPE, FUSED>(rows, impactHdr, plan, motion, mass, ctx, tape, bar, phase, cnt, grid > 0);
    }
}

template <int MODE, bool MADE, bool TAPE>
__global__ void __launch_bounds__(kPassThreads) k_pass(Rows rows, const uint4* impactHdr, PassPlan plan,
                                                       float4* motion, const float4* mass, Ctx ctx,
                                                       BarrierLines* bar, BarrierBase base, DevCounts* cnt,
                                                       Tape tape) {
    BarrierPhase phase(base);
    passRun<MODE, MADE, TAPE, false>(rows, impactHdr, plan, motion, mass, ctx, bar, phase, cnt, tape);
}

// The barriers a pass launch of `plan` arrives at.
__host__ __device__ __forceinline__ uint32_t passBarriers(const PassPlan& plan) {
    if (plan.n == 0) { return 0; }
    return (plan.tailFrom == plan.n ? plan.n - 1 : plan.tailFrom) + (plan.sepRows > 0 ? 1u : 0u);
}

// ---- warm start, one body at a time ----------------------------------------------------------------
//
// warmStartBatches applies each coloured row's stored impulses to its moving bodies, row by row
// (src/lanes.cpp:195-217, 413-417). A warm-start push does not depend on the velocities it changes,
// so each body can apply its own rows' pushes, in row order, by itself: one launch instead of one
// per colour, with the same operations on each body in the same order.

// A body's entry for one of its rows: the row and the body's side (entries sort by row). The row's
// point count is read from its header: a refresh can change it while the lists stay.
__device__ __forceinline__ uint32_t warmEntry(uint32_t r, bool sideB) { return (r << 1) | (sideB ? 1u : 0u); }
__device__ __forceinline__ uint32_t entryRow(uint32_t e) { return e >> 1; }

__global__ void k_bep_count(uint32_t n, const uint4* hdr, uint32_t* deg) {
    for (uint32_t r = gid(); r < n; r += gstride()) {
        const uint4 h = hdr[r];
        if ((h.z & kCountMask) == 0) { continue; }
        if (h.z & kMoveA) { atomicAdd(&deg[h.x], 1u); }
        if (h.z & kMoveB) { atomicAdd(&deg[h.y], 1u); }
    }
}

__global__ void k_bep_fill(uint32_t n, const uint4* hdr, const uint32_t* start, uint32_t* cursor, uint32_t* ep) {
    for (uint32_t r = gid(); r < n; r += gstride()) {
        const uint4 h = hdr[r];
        if ((h.z & kCountMask) == 0) { continue; }
        const uint32_t tag = warmEntry(r, false);
        if (h.z & kMoveA) { ep[start[h.x] + atomicAdd(&cursor[h.x], 1u)] = tag; }
        if (h.z & kMoveB) { ep[start[h.y] + atomicAdd(&cursor[h.y], 1u)] = tag | 1u; }
    }
}

__global__ void k_bep_sort(uint32_t nb, const uint32_t* deg, const uint32_t* start, uint32_t* ep) {
    for (uint32_t x = gid(); x < nb; x += gstride()) {
        uint32_t* e = ep + start[x];
        const uint32_t d = deg[x];
        for (uint32_t i = 1; i < d; ++i) {
            const uint32_t v = e[i];
            uint32_t j = i;
            while (j > 0 && e[j - 1] > v) {
                e[j] = e[j - 1];
                --j;
            }
            e[j] = v;
        }
    }
}

// One row's warm-start data, this body's side only: impulses, tangents and two float4 per point and
// axis.
struct WarmSide {
    int count;
    float4 t1, t2;
    float4 im[4];
    float4 q[4][3][2];
};

__device__ __forceinline__ void warmLoad(const Rows& rows, uint32_t e, WarmSide& w) {
    const uint32_t r = entryRow(e);
    const int count = (int)(rows.hdr[r].z & kCountMask);
    const uint32_t side = e & 1u;
    w.count = count;
    w.t1 = rows.tg1[r];
    w.t2 = rows.tg2[r];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i < count) {
            w.im[i] = rows.imp[(uint32_t)i * rows.cap + r];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float4* base = rows.ax + (size_t)((i * 3 + a) * 5 + side) * rows.cap + r;
                w.q[i][a][0] = base[0];
                w.q[i][a][1] = base[rows.cap];
            }
        }
    }
}

// warmStartLanes' pushes of one row on this body (src/lanes.cpp:195-217), in point and axis order.
__device__ __forceinline__ void warmApply(uint32_t e, const WarmSide& w, Motion& m) {
    const int count = w.count;
    const bool sideB = e & 1u;
    const V3 dir1 = vec(w.t1.x, w.t1.y, w.t1.z), dir2 = vec(w.t2.x, w.t2.y, w.t2.z);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (i >= count) { continue; }
        const V3 ti = vec(w.im[i].y, w.im[i].z, w.im[i].w);
        const float lambda[3] = { w.im[i].x, dot(ti, dir1), dot(ti, dir2) };
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float4 q0 = w.q[i][a][0], q1 = w.q[i][a][1];
            if (sideB) {
                m.v = m.v + vec(q0.z, q0.w, q1.x) * lambda[a];
                m.w = m.w + vec(q1.y, q1.z, q1.w) * lambda[a];
            } else {
                m.v = m.v - vec(q0.x, q0.y, q0.z) * lambda[a];
                m.w = m.w - vec(q0.w, q1.x, q1.y) * lambda[a];
            }
        }
    }
}

// Each body applies its rows in row order; a row's loads go out while the row before it is applied
// (two rows in flight).
__device__ __forceinline__ void warmBodyRows(uint32_t d, const uint32_t* __restrict__ list, const Rows& rows, Motion& m) {
    {
        WarmSide a, b;
        uint32_t ea = list[0], eb = d > 1 ? list[1] : 0u;
        warmLoad(rows, ea, a);
        for (uint32_t k = 0; k < d; k += 2) {
            if (k + 1 < d) { warmLoad(rows, eb, b); }
            const uint32_t ec = k + 2 < d ? list[k + 2] : 0u;
            warmApply(ea, a, m);
            if (k + 2 < d) { warmLoad(rows, ec, a); }
            const uint32_t ed = k + 3 < d ? list[k + 3] : 0u;
            if (k + 1 < d) { warmApply(eb, b, m); }
            ea = ec;
            eb = ed;
        }
    }
}

// ---- touch events: reportTouches (src/world.cpp:1424-1468) -------------------------------------------
//
// One thread per pair: the thread of a pair's first contact (a contact whose key differs from the one
// before it) walks the pair's contacts in order, as the reference's loop does, decides whether the pair
// touches and which pushed point approached fastest, flags an event when the pair's state changed
// (and writes the new state to all its contacts) and a hit when a point was fast enough. Every thread
// counts its own contact's touching() (src/internal.h:497-502). Flags are compacted by a scan.
__global__ void k_touches(tfs_contact* contacts, uint32_t nc, const tfs_body* bodies, float hitThreshold, float linearSlop,
                          uint32_t* touchFlag, uint32_t* hitFlag, tfs_hit* hitData, uint32_t* touching) {
    uint32_t count = 0;
    for (uint32_t i = gid(); i < nc; i += gstride()) {
        const tfs_contact& self = contacts[i];
        bool overlapping = false;
        TF_POINTS(k, self.count) {
            overlapping = overlapping || self.points[k].separation < 0.0f || self.points[k].normal_impulse > 0.0f;
        }
        count += overlapping ? 1u : 0u;
        uint32_t ev = 0, hit = 0;
        if (i == 0 || contacts[i - 1].key != self.key) {
            const bool was = self.touched != 0;
            const float apart = was ? linearSlop : 0.0f;
            bool touch = false;
            float fastest = hitThreshold;
            uint32_t struck = kNoSlot;
            int point = 0;
            uint32_t last = i;
            for (; last < nc && contacts[last].key == self.key; ++last) {
                const tfs_contact& c = contacts[last];
                const tfs_body& a = bodies[c.body_a];
                const tfs_body& b = bodies[c.body_b];
                const bool solved = movingBody(a) || movingBody(b);
                for (int k = 0; k < c.count && k < 4; ++k) {
                    const tfs_point& p = c.points[k];
                    if (p.max_normal_impulse > 0.0f) {
                        touch = true;
                        if (-p.relative_velocity > fastest && solved) {
                            fastest = -p.relative_velocity;
                            struck = last;
                            point = k;
                        }
                    } else if (p.separation < apart) {
                        touch = true;
                    }
                }
            }
            if (touch != was) {
                ev = 1;
                for (uint32_t k = i; k < last; ++k) { contacts[k].touched = touch ? 1 : 0; }
            }
            if (struck != kNoSlot) {
                hit = 1;
                const tfs_contact& c = contacts[struck];
                const tfs_body& a = bodies[c.body_a];
                const V3 at = ld3(a.p) + rotate(ld4(a.q), ld3(c.points[point].local_a));
                tfs_hit h;
                h.contact = struck;
                h.point = (uint32_t)point;
                st3(h.position, at);
                for (int d = 0; d < 3; ++d) { h.normal[d] = c.normal[d]; }
                h.speed = fastest;
                hitData[i] = h;
            }
        }
        touchFlag[i] = ev;
        hitFlag[i] = hit;
    }
    waveAdd(touching, count);
}

// Each event gets its pair's new state: (first contact << 1) | began.
__global__ void k_touch_events(uint32_t* list, const uint32_t* nPtr, const tfs_contact* contacts) {
    const uint32_t n = *nPtr;
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const uint32_t i = list[k];
        list[k] = (i << 1) | (contacts[i].touched ? 1u : 0u);
    }
}

__global__ void k_gather_hits(const uint32_t* list, const uint32_t* nPtr, const tfs_hit* hitData, tfs_hit* out) {
    const uint32_t n = *nPtr;
    for (uint32_t k = gid(); k < n; k += gstride()) { out[k] = hitData[list[k]]; }
}

// ---- the contact stream: updateContacts' solver half (src/world.cpp:880-989) -----------------------
//
// Pairs and old contacts are both in key order. Per pair: its old contacts by binary search, whether
// they are all kept, and how many contacts it leaves; per old contact: whether its pair rests (kept
// as it is). Scans place both lists, and the two merge by key: a pair's outputs go after the resting
// contacts with smaller keys, and a resting contact after the outputs of the pairs with smaller keys
// (the two key sets are disjoint: a pair has a body dynamic and awake or a moving kinematic one, a
// resting contact has none).

__device__ __forceinline__ uint32_t lowerKey(const tfs_contact* c, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (c[mid].key < key) { lo = mid + 1; } else { hi = mid; }
    }
    return lo;
}

__global__ void k_stream_match(const tfs_pair* pairs, uint32_t np, const tfs_contact* old, const tfs_contact_ext* oldExt,
                               uint32_t nOld, const tfs_body* bodies, float reach, float reuseCos, uint32_t* lo, uint32_t* hi,
                               uint32_t* keep, uint32_t* count) {
    for (uint32_t p = gid(); p < np; p += gstride()) {
        const tfs_pair& x = pairs[p];
        const uint32_t l = lowerKey(old, nOld, x.key);
        uint32_t h = l;
        while (h < nOld && old[h].key == x.key) { ++h; }
        const tfs_body& a = bodies[x.body_a];
        const tfs_body& b = bodies[x.body_b];
        bool k = h > l;
        for (uint32_t i = l; i < h && k; ++i) { k = stillFits(oldExt[i], a, b, reach, reuseCos); }
        lo[p] = l;
        hi[p] = h;
        keep[p] = k ? 1u : 0u;
        count[p] = k ? h - l : x.count;
    }
}

// src/world.cpp:972-984: an old contact whose bodies are neither dynamic and awake nor moving kinematic,
// one of them dynamic, stays as it is.
__global__ void k_stream_resting(const tfs_contact* old, uint32_t nOld, const tfs_body* bodies, uint32_t* flag) {
    for (uint32_t i = gid(); i < nOld; i += gstride()) {
        const tfs_body& a = bodies[old[i].body_a];
        const tfs_body& b = bodies[old[i].body_b];
        const bool rest = !movingAwake(a, ld3(a.v), ld3(a.w)) && !movingAwake(b, ld3(b.v), ld3(b.w)) &&
                          (a.type == TFS_DYNAMIC || b.type == TFS_DYNAMIC);
        flag[i] = rest ? 1u : 0u;
    }
}

// Resting contacts (listed) with keys below `key`.
__device__ __forceinline__ uint32_t restingBelow(const tfs_contact* old, const uint32_t* list, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (old[list[mid]].key < key) { lo = mid + 1; } else { hi = mid; }
    }
    return lo;
}

__global__ void k_stream_fill(const tfs_pair* pairs, uint32_t np, const uint32_t* lo, const uint32_t* hi, const uint32_t* keep,
                              const uint32_t* start, const tfs_manifold* mf, const tfs_contact* old,
                              const tfs_contact_ext* oldExt, const uint32_t* restList, const uint32_t* nRestPtr,
                              const tfs_body* bodies, float speculative, float matchSq, tfs_contact* out,
                              tfs_contact_ext* outExt) {
    const uint32_t nRest = *nRestPtr;
    for (uint32_t p = gid(); p < np; p += gstride()) {
        const tfs_pair& x = pairs[p];
        const uint32_t base = start[p] + restingBelow(old, restList, nRest, x.key);
        const tfs_body& a = bodies[x.body_a];
        const tfs_body& b = bodies[x.body_b];
        const uint32_t l = lo[p], h = hi[p];
        uint32_t made = 0;
        if (keep[p]) {
            for (uint32_t i = l; i < h; ++i, ++made) {
                tfs_contact c = old[i];
                const tfs_contact_ext e = oldExt[i];
                measureAgain(a, b, c, e, speculative);
                c.friction = x.friction;
                c.restitution = x.restitution;
                c.rolling_resistance = x.rolling_resistance;
                out[base + made] = c;
                outExt[base + made] = e;
            }
            continue;
        }
        const uint8_t touched = h > l ? old[l].touched : 0;
        for (uint32_t j = 0; j < x.count; ++j, ++made) {
            tfs_contact c;
            memset(&c, 0, sizeof c);
            tfs_contact_ext e;
            memset(&e, 0, sizeof e);
            c.key = x.key;
            c.body_a = x.body_a;
            c.body_b = x.body_b;
            c.touched = touched;
            e.part = j;
            contactFromManifold(a, b, mf[x.first + j], c, e, speculative);
            // closestNormal (src/world.cpp:543-549): the old contact with points whose normal is closest
            uint32_t best = kNoSlot;
            const V3 n = ld3(c.normal);
            for (uint32_t i = l; i < h; ++i) {
                if (old[i].count > 0 && (best == kNoSlot || dot(ld3(old[i].normal), n) > dot(ld3(old[best].normal), n))) {
                    best = i;
                }
            }
            if (best != kNoSlot) { carryImpulses(old[best], c, matchSq); }
            c.friction = x.friction;
            c.restitution = x.restitution;
            c.rolling_resistance = x.rolling_resistance;
            out[base + made] = c;
            outExt[base + made] = e;
        }
    }
}

// Pairs with keys below `key`.
__device__ __forceinline__ uint32_t pairsBelow(const tfs_pair* pairs, uint32_t n, uint64_t key) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (pairs[mid].key < key) { lo = mid + 1; } else { hi = mid; }
    }
    return lo;
}

__global__ void k_stream_rest_fill(const uint32_t* restList, const uint32_t* nRestPtr, const tfs_contact* old,
                                   const tfs_contact_ext* oldExt, const tfs_pair* pairs, uint32_t np, const uint32_t* start,
                                   const uint32_t* pairTotal, tfs_contact* out, tfs_contact_ext* outExt) {
    const uint32_t nRest = *nRestPtr;
    for (uint32_t k = gid(); k < nRest; k += gstride()) {
        const uint32_t i = restList[k];
        const uint32_t p = pairsBelow(pairs, np, old[i].key);
        const uint32_t before = p < np ? start[p] : *pairTotal;
        out[k + before] = old[i];
        outExt[k + before] = oldExt[i];
    }
}

// ---- the fixed-contact model and its adjoint ---------------------------------------------------------
//
// The solve block run for several steps over one contact set, fixed at the first step: no
// narrowphase between steps, the contact frames, arms and effective masses (the packed rows) kept,
// and each point's adjusted separation rebased by the bodies' travel along the normal between
// steps (rebaseContact, src/solver.cpp:293-306, with dq = identity: rotations are left out of the
// point positions). Per substep: integrateVelocities, the warm start, the biased pass, the
// positions (com and dp only), the relax pass (src/world.cpp:1671-1681); no impact pass and no
// restitution. The adjoint runs the reverse sweep of each step from a checkpoint of the step's start,
// solving the step again to record what each row saw (Tape).

__device__ __forceinline__ bool movesFreely(const tfs_body& b) {
    return b.alive && b.type != TFS_STATIC && !(b.type == TFS_DYNAMIC && !b.awake);
}

// integratePosition's centre of mass and travel (src/world.cpp:1161-1172), without the rotation.
__global__ void k_fx_positions(tfs_body* bodies, uint32_t n, float4* motion, float h) {
    for (uint32_t i = gid(); i < n; i += gstride()) {
        tfs_body& b = bodies[i];
        if (!movesFreely(b)) { continue; }
        const float4 v = motion[4 * i], dp = motion[4 * i + 2];
        const V3 vh = vec(v.x, v.y, v.z) * h;
        const V3 com = ld3(b.com) + vh;
        st3(b.com, com);
        const V3 d = vec(dp.x, dp.y, dp.z) + vh;
        motion[4 * i + 2] = make_float4(d.x, d.y, d.z, 0.0f);
    }
}

// rebaseContact's separation shift for every coloured row (src/solver.cpp:293-306).
__global__ void k_fx_rebase(Rows rows, uint32_t n, const float4* motion) {
    for (uint32_t r = gid(); r < n; r += gstride()) {
        const uint4 h = rows.hdr[r];
        const int count = (int)(h.z & kCountMask);
        const float4 pa = motion[4 * h.x + 2], pb = motion[4 * h.y + 2], nr = rows.nrm[r];
        const float shift = dot(vec(pb.x, pb.y, pb.z) - vec(pa.x, pa.y, pa.z), vec(nr.x, nr.y, nr.z));
        TF_POINTS(k, count) { rows.pa[(uint32_t)k * rows.cap + r].w += shift; }
    }
}

__global__ void k_fx_reset_dp(uint32_t n, float4* motion) {
    for (uint32_t i = gid(); i < n; i += gstride()) { motion[4 * i + 2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f); }
}

// A step's checkpoint: per body its velocities and centre of mass (3 float4), per row its points'
// impulses and adjusted separations (4 + 1 float4).
struct Checkpoint {
    float4* body;   // [3 * i + q]
    float4* row;    // [q * cap + r], q = 0..3 impulses, 4 separations
};

__global__ void k_fx_save(const tfs_body* bodies, uint32_t nb, const float4* motion, Rows rows, uint32_t nr,
                          Checkpoint cp) {
    for (uint32_t i = gid(); i < nb; i += gstride()) {
        cp.body[3 * i] = motion[4 * i];
        cp.body[3 * i + 1] = motion[4 * i + 1];
        const V3 c = ld3(bodies[i].com);
        cp.body[3 * i + 2] = make_float4(c.x, c.y, c.z, 0.0f);
    }
    const uint32_t cap = rows.cap;
    for (uint32_t r = gid(); r < nr; r += gstride()) {
        float4 sep = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        float* sp = &sep.x;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            cp.row[(uint32_t)k * cap + r] = rows.imp[(uint32_t)k * cap + r];
            sp[k] = rows.pa[(uint32_t)k * cap + r].w;
        }
        cp.row[4u * cap + r] = sep;
    }
}

__global__ void k_fx_restore(tfs_body* bodies, uint32_t nb, float4* motion, Rows rows, uint32_t nr, Checkpoint cp) {
    for (uint32_t i = gid(); i < nb; i += gstride()) {
        motion[4 * i] = cp.body[3 * i];
        motion[4 * i + 1] = cp.body[3 * i + 1];
        motion[4 * i + 2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float4 c = cp.body[3 * i + 2];
        st3(bodies[i].com, vec(c.x, c.y, c.z));
    }
    const uint32_t cap = rows.cap;
    for (uint32_t r = gid(); r < nr; r += gstride()) {
        const float4 sep = cp.row[4u * cap + r];
        const float* sp = &sep.x;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            rows.imp[(uint32_t)k * cap + r] = cp.row[(uint32_t)k * cap + r];
            rows.pa[(uint32_t)k * cap + r].w = sp[k];
        }
    }
}

// The adjoint state: per body d loss / d (v, w) (bar[2i], bar[2i + 1]), d loss / d dp and the
// fixed d loss / d com of the last step; per row point d loss / d (normal impulse, tangent impulse)
// and d loss / d adjusted separation; per row d loss / d friction.
struct Adjoint {
    float4* bar;
    float4* dpBar;
    float4* comBar;
    float4* impBar;   // [k * cap + r]
    float* sepBar;    // [k * cap + r]
    float* muBar;     // [r]
};

__global__ void k_adj_seed(uint32_t nb, const float* dlcom, uint32_t nr, uint32_t cap, Adjoint g) {
    for (uint32_t i = gid(); i < nb; i += gstride()) {
        g.comBar[i] = make_float4(dlcom[3 * i], dlcom[3 * i + 1], dlcom[3 * i + 2], 0.0f);
        g.bar[2 * i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        g.bar[2 * i + 1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        g.dpBar[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    for (uint32_t r = gid(); r < nr; r += gstride()) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            g.impBar[(uint32_t)k * cap + r] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            g.sepBar[(uint32_t)k * cap + r] = 0.0f;
        }
        g.muBar[r] = 0.0f;
    }
}

// The reverse of the step's separation rebase: each moving body's d loss / d dp at the end of the
// step gathers +-n times its rows' separation adjoints (+ on side B, - on side A), in row order.
__global__ void k_adj_rebase(uint32_t nb, const uint32_t* deg, const uint32_t* start, const uint32_t* ep, Rows rows,
                             Adjoint g) {
    for (uint32_t x = gid(); x < nb; x += gstride()) {
        const uint32_t d = deg[x];
        const uint32_t* list = ep + start[x];
        V3 sum = vec(0.0f, 0.0f, 0.0f);
        for (uint32_t k = 0; k < d; ++k) {
            const uint32_t r = entryRow(list[k]);
            const int count = (int)(rows.hdr[r].z & kCountMask);
            float s = 0.0f;
            TF_POINTS(q, count) { s += g.sepBar[(uint32_t)q * rows.cap + r]; }
            const float4 nr = rows.nrm[r];
            const V3 t = vec(nr.x, nr.y, nr.z) * s;
            sum = (list[k] & 1u) ? sum + t : sum - t;
        }
        g.dpBar[x] = make_float4(sum.x, sum.y, sum.z, 0.0f);
    }
}

// The reverse of k_fx_positions: com += v h and dp += v h give v their adjoints times h.
__global__ void k_adj_positions(const tfs_body* bodies, uint32_t n, float h, Adjoint g) {
    for (uint32_t i = gid(); i < n; i += gstride()) {
        if (!movesFreely(bodies[i])) { continue; }
        const float4 v = g.bar[2 * i], c = g.comBar[i], d = g.dpBar[i];
        const V3 x = vec(v.x, v.y, v.z) + (vec(c.x, c.y, c.z) + vec(d.x, d.y, d.z)) * h;
        g.bar[2 * i] = make_float4(x.x, x.y, x.z, 0.0f);
    }
}

// The reverse of the warm start: the pushes do not depend on the velocities, so the velocity
// adjoints pass through and each stored impulse takes its push's adjoint (pushBar).
__global__ void __launch_bounds__(kThreads) k_adj_warm(Rows rows, uint32_t n, Adjoint g) {
    for (uint32_t r = gid(); r < n; r += gstride()) {
        const uint4 h = rows.hdr[r];
        const int count = (int)(h.z & kCountMask);
        if (count == 0) { continue; }
        const bool moveA = h.z & kMoveA, moveB = h.z & kMoveB;
        const float4 a0 = g.bar[2 * h.x], a1 = g.bar[2 * h.x + 1], b0 = g.bar[2 * h.y], b1 = g.bar[2 * h.y + 1];
        const MotionBar mb{ vec(a0.x, a0.y, a0.z), vec(a1.x, a1.y, a1.z), vec(b0.x, b0.y, b0.z), vec(b1.x, b1.y, b1.z) };
        const float4 t1 = rows.tg1[r], t2 = rows.tg2[r];
        const V3 d1 = vec(t1.x, t1.y, t1.z), d2 = vec(t2.x, t2.y, t2.z);
        const LoadedAxes ax{ rows, r };
        TF_POINTS(k, count) {
            const uint32_t j = (uint32_t)k * rows.cap + r;
            const float4 ib = g.impBar[j];
            const float n0 = pushBar(ax(k, 0), mb, moveA, moveB);
            const float n1 = pushBar(ax(k, 1), mb, moveA, moveB);
            const float n2 = pushBar(ax(k, 2), mb, moveA, moveB);
            const V3 t = vec(ib.y, ib.z, ib.w) + (d1 * n1 + d2 * n2);
            g.impBar[j] = make_float4(ib.x + n0, t.x, t.y, t.z);
        }
    }
}

// The reverse of one row of a pass, from what the recording pass saw (Tape).
template <bool BIAS, bool COHERENT, bool MADE>
__device__ __forceinline__ void rowAdjoint(const Rows& rows, uint32_t r, const Tape& tape, const float4* mass,
                                           const Ctx& ctx, const Adjoint& g) {
    using X = Access<COHERENT>;
    const uint4 h = rows.hdr[r];
    const int count = (int)(h.z & kCountMask);
    if (count == 0) { return; }
    const float4 n = rows.nrm[r], t1 = rows.tg1[r], t2 = rows.tg2[r];
    const Row row{ h.z, vec(n.x, n.y, n.z), vec(t1.x, t1.y, t1.z), vec(t2.x, t2.y, t2.z), n.w, t1.w };
    const uint32_t c = tape.cap;
    RowPoint p[4];
    float niB[4] = { 0.0f, 0.0f, 0.0f, 0.0f }, sepB[4] = { 0.0f, 0.0f, 0.0f, 0.0f };
    V3 tiB[4];
    float4 axv[4][MADE ? 1 : 15];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < count) {
            const uint32_t j = (uint32_t)k * rows.cap + r;
            const float4 im = tape.imp[(uint32_t)k * c + r], ib = g.impBar[j];
            if (MADE) {
                const float4 a = rows.pa[j], b = rows.pb[j];
                p[k].ra = vec(a.x, a.y, a.z);
                p[k].rb = vec(b.x, b.y, b.z);
            }
            p[k].ni = im.x;
            p[k].ti = vec(im.y, im.z, im.w);
            p[k].nimax = 0.0f;
            niB[k] = ib.x;
            tiB[k] = vec(ib.y, ib.z, ib.w);
            if (!MADE) {
#pragma unroll
                for (int q = 0; q < 15; ++q) { axv[k][q] = rows.ax[(size_t)(k * 15 + q) * rows.cap + r]; }
            }
        }
    }
    const float4 sp4 = tape.sep[r];
    const float sep[4] = { sp4.x, sp4.y, sp4.z, sp4.w };
    const float4 va = tape.v[r], wa = tape.v[c + r], vb = tape.v[2 * c + r], wb = tape.v[3 * c + r];
    const Q4 one = quat(0.0f, 0.0f, 0.0f, 1.0f);
    const V3 zero = vec(0.0f, 0.0f, 0.0f);
    const Motion A{ vec(va.x, va.y, va.z), vec(wa.x, wa.y, wa.z), zero, one };
    const Motion B{ vec(vb.x, vb.y, vb.z), vec(wb.x, wb.y, wb.z), zero, one };
    const float4 ga = X::ld(&g.bar[2 * h.x]), gwa = X::ld(&g.bar[2 * h.x + 1]);
    const float4 gb = X::ld(&g.bar[2 * h.y]), gwb = X::ld(&g.bar[2 * h.y + 1]);
    MotionBar mb{ vec(ga.x, ga.y, ga.z), vec(gwa.x, gwa.y, gwa.z), vec(gb.x, gb.y, gb.z), vec(gwb.x, gwb.y, gwb.z) };
    float muB = 0.0f;
    if (MADE) {
        const RowMadeAxes ax{ loadMass(mass, h.x), loadMass(mass, h.y), row, p };
        solveRowAdjoint<BIAS>(row, p, sep, ax, A, B, ctx, count, mb, niB, tiB, sepB, muB);
    } else {
        const RegAxes<MADE ? 1 : 15> ax{ axv };
        solveRowAdjoint<BIAS>(row, p, sep, ax, A, B, ctx, count, mb, niB, tiB, sepB, muB);
    }
    float sepSum = 0.0f;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (k < count) {
            const uint32_t j = (uint32_t)k * rows.cap + r;
            g.impBar[j] = make_float4(niB[k], tiB[k].x, tiB[k].y, tiB[k].z);
            g.sepBar[j] += sepB[k];
            sepSum += sepB[k];
        }
    }
    g.muBar[r] += muB;
    const V3 dn = row.n * sepSum;
    if (h.z & kMoveA) {
        X::st(&g.bar[2 * h.x], make_float4(mb.va.x, mb.va.y, mb.va.z, 0.0f));
        X::st(&g.bar[2 * h.x + 1], make_float4(mb.wa.x, mb.wa.y, mb.wa.z, 0.0f));
        const float4 d = X::ld(&g.dpBar[h.x]);
        const V3 x = vec(d.x, d.y, d.z) - dn;
        X::st(&g.dpBar[h.x], make_float4(x.x, x.y, x.z, 0.0f));
    }
    if (h.z & kMoveB) {
        X::st(&g.bar[2 * h.y], make_float4(mb.vb.x, mb.vb.y, mb.vb.z, 0.0f));
        X::st(&g.bar[2 * h.y + 1], make_float4(mb.wb.x, mb.wb.y, mb.wb.z, 0.0f));
        const float4 d = X::ld(&g.dpBar[h.y]);
        const V3 x = vec(d.x, d.y, d.z) + dn;
        X::st(&g.dpBar[h.y], make_float4(x.x, x.y, x.z, 0.0f));
    }
}

// The reverse of a pass in one launch with a grid barrier: the plan lists the colours last to first.
template <bool BIAS, bool MADE>
__global__ void __launch_bounds__(kPassThreads) k_pass_adj(Rows rows, PassPlan plan, Tape tape, const float4* mass,
                                                           Ctx ctx, Adjoint g, BarrierLines* bar, BarrierBase base,
                                                           DevCounts* cnt) {
    BarrierPhase phase(base);
    const uint32_t slot = passSlot(), stride = gstride();
    for (uint32_t s = 0; s < plan.n; ++s) {
        for (uint32_t r = plan.begin[s] + slot; r < plan.end[s]; r += stride) {
            rowAdjoint<BIAS, true, MADE>(rows, r, tape, mass, ctx, g);
        }
        if (s + 1 < plan.n) { gridSync(bar, phase, cnt); }
    }
}

// Per row, d loss / d friction goes to its contact.
__global__ void k_adj_friction(Rows rows, uint32_t n, const float* muBar, float* perContact) {
    for (uint32_t r = gid(); r < n; r += gstride()) { perContact[rows.hdr[r].w] = muBar[r]; }
}

// ---- the rest list in levels ------------------------------------------------------------------------
//
// Contacts outside the colours are solved one at a time in list order (solveContacts, warmStart,
// applyRestitution: src/solver.cpp:319-460). A contact changes the velocities of its moving bodies
// only, except that rolling resistance (resistTurning) writes both bodies whatever they are. So a
// body is tracked when it moves or a rolling contact of the list touches it, and two contacts that
// share no tracked body commute. Each contact's level is 1 + the highest level among the earlier
// contacts of the list that share a tracked body with it; solving level after level, the contacts
// of one level in parallel, gives every tracked body the same updates in the same order as the list
// does, to the bit.

struct RestScratch {
    uint8_t* tracked;
    uint32_t *deg, *start, *cursor, *ep, *rank, *level, *count;
    unsigned long long* word;
};

__device__ __forceinline__ bool rollingContact(const tfs_contact& c) { return c.rolling_resistance > 0.0f; }

__global__ void k_rest_track(uint32_t nb, const uint8_t* moving, RestScratch z) {
    for (uint32_t i = gid(); i < nb; i += gstride()) {
        z.tracked[i] = moving[i];
        z.deg[i] = z.cursor[i] = 0;
        z.word[i] = 0ull;
    }
}

__global__ void k_rest_roll(const uint32_t* list, uint32_t n, const tfs_contact* contacts, RestScratch z) {
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const tfs_contact& c = contacts[list[k]];
        if (rollingContact(c)) { z.tracked[c.body_a] = z.tracked[c.body_b] = 1; }
        z.count[k] = 0;
    }
}

__global__ void k_rest_ep_count(const uint32_t* list, uint32_t n, const tfs_contact* contacts, RestScratch z) {
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const tfs_contact& c = contacts[list[k]];
        if (z.tracked[c.body_a]) { atomicAdd(&z.deg[c.body_a], 1u); }
        if (z.tracked[c.body_b]) { atomicAdd(&z.deg[c.body_b], 1u); }
    }
}

__global__ void k_rest_ep_fill(const uint32_t* list, uint32_t n, const tfs_contact* contacts, RestScratch z) {
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const tfs_contact& c = contacts[list[k]];
        if (z.tracked[c.body_a]) { z.ep[z.start[c.body_a] + atomicAdd(&z.cursor[c.body_a], 1u)] = k << 1; }
        if (z.tracked[c.body_b]) { z.ep[z.start[c.body_b] + atomicAdd(&z.cursor[c.body_b], 1u)] = (k << 1) | 1u; }
    }
}

// The first level of a rest list that fits one rest work-group (levels before it are wide: they run
// over many blocks). Written as ~level with atomicMax, so the counters' zero start means "none".
__global__ void k_rest_narrow(const uint32_t* levelStart, const uint32_t* levels, uint32_t width, uint32_t* first) {
    const uint32_t n = *levels;
    for (uint32_t lv = gid(); lv < n; lv += gstride()) {
        if (levelStart[lv + 1] - levelStart[lv] <= width) { atomicMax(first, ~lv); }
    }
}

// The levels, as the colouring's dataflow: a contact proceeds once its tracked bodies have seen all
// their earlier contacts; each body's word is (contacts seen << 32) | its highest level so far.
// As k_color: 1,024-contact chunks by ticket, the links from earlier chunks through the bodies' global
// words (a wave waits for all of its lanes' before the next part), the links within the chunk through
// LDS slots, each finished contact leaving its bodies' words in its slot.
__global__ void __launch_bounds__(kColorThreads)
    k_rest_level(const uint32_t* list, uint32_t n, const tfs_contact* contacts, RestScratch z, uint32_t* ticket,
                 uint32_t* levels, DevCounts* cnt) {
    __shared__ uint32_t chunk;
    __shared__ uint32_t slotDone[kColorThreads];
    __shared__ unsigned long long slotWord[kColorThreads][2];
    for (;;) {
        if (threadIdx.x == 0) { chunk = atomicAdd(ticket, 1u); }
        slotDone[threadIdx.x] = 0u;
        __syncthreads();
        const uint32_t base = chunk * blockDim.x;
        const uint32_t k = base + threadIdx.x;
        if (base >= n) { break; }
        if (k < n) {
            const tfs_contact& c = contacts[list[k]];
            const uint32_t a = c.body_a, b = c.body_b;
            const bool ta = z.tracked[a], tb = z.tracked[b];
            const uint32_t ra = ta ? z.rank[k << 1] : 0u, rb = tb ? z.rank[(k << 1) | 1u] : 0u;
            uint32_t la = 0xffffffffu, lb = 0xffffffffu;
            if (ta && ra > 0) {
                const uint32_t e = z.ep[z.start[a] + ra - 1];
                if ((e >> 1) >= base) { la = (((e >> 1) - base) << 1) | (e & 1u); }
            }
            if (tb && rb > 0) {
                const uint32_t e = z.ep[z.start[b] + rb - 1];
                if ((e >> 1) >= base) { lb = (((e >> 1) - base) << 1) | (e & 1u); }
            }
            unsigned long long wa = 0ull, wb = 0ull;
            bool readyA = !ta || la != 0xffffffffu, readyB = !tb || lb != 0xffffffffu;
            uint32_t spins = 0;
            while (!(readyA && readyB)) {
                if (!readyA) {
                    wa = loadWord(&z.word[a]);
                    readyA = (uint32_t)(wa >> 32) == ra;
                }
                if (!readyB) {
                    wb = loadWord(&z.word[b]);
                    readyB = (uint32_t)(wb >> 32) == rb;
                }
                if (readyA && readyB) { break; }
                if (++spins > kSpinLimit) {
                    atomicOr(&cnt->error, 8u);
                    readyA = readyB = true;
                    la = lb = 0xffffffffu;
                } else {
                    __builtin_amdgcn_s_sleep(1);
                }
            }
            bool done = false;
            while (!done) {
                bool okA = true, okB = true;
                if (ta && la != 0xffffffffu) {
                    okA = __hip_atomic_load(&slotDone[la >> 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u;
                    if (okA) { wa = slotWord[la >> 1][la & 1u]; }
                }
                if (tb && lb != 0xffffffffu) {
                    okB = __hip_atomic_load(&slotDone[lb >> 1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) != 0u;
                    if (okB) { wb = slotWord[lb >> 1][lb & 1u]; }
                }
                if (okA && okB) {
                    const uint32_t lv = 1u + max((uint32_t)wa, (uint32_t)wb);
                    const unsigned long long na = ((unsigned long long)(ra + 1) << 32) | lv;
                    const unsigned long long nb = ((unsigned long long)(rb + 1) << 32) | lv;
                    if (ta) { storeWord(&z.word[a], na); }
                    if (tb) { storeWord(&z.word[b], nb); }
                    slotWord[threadIdx.x][0] = na;
                    slotWord[threadIdx.x][1] = nb;
                    __hip_atomic_store(&slotDone[threadIdx.x], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    z.level[k] = lv - 1u;
                    atomicAdd(&z.count[lv - 1u], 1u);
                    atomicMax(levels, lv);
                    done = true;
                } else if (++spins > kSpinLimit) {
                    atomicOr(&cnt->error, 8u);
                    __hip_atomic_store(&slotDone[threadIdx.x], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
                    z.level[k] = 0u;
                    done = true;
                } else {
                    __builtin_amdgcn_s_sleep(1);
                }
            }
        }
        __syncthreads();
    }
}

// Contacts of the list grouped by level (any order within a level: its contacts commute), each entry
// (contact, body a, body b, which bodies it writes: 1 a, 2 b).
__global__ void k_rest_place(const uint32_t* list, uint32_t n, const tfs_contact* contacts, const uint8_t* moving,
                             RestScratch z, const uint32_t* levelStart, uint4* entries, uint32_t* shape) {
    uint32_t sum[4] = { 0u, 0u, 0u, 0u };
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const uint32_t lv = z.level[k];
        const uint32_t ci = list[k];
        const tfs_contact& c = contacts[ci];
        const bool rolling = rollingContact(c);
        const uint32_t writes = (rolling || moving[c.body_a] ? 1u : 0u) | (rolling || moving[c.body_b] ? 2u : 0u) |
                                (moving[c.body_a] ? 0u : 4u) | (moving[c.body_b] ? 0u : 8u);
        entries[levelStart[lv] + atomicSub(&z.count[lv], 1u) - 1u] = make_uint4(ci, c.body_a, c.body_b, writes);
        sum[c.planar ? 0 : 2] += 1u;
        sum[c.planar ? 1 : 3] += (uint32_t)c.count;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) { waveAdd(&shape[q], sum[q]); }
}

// The active rest contacts, packed in the level order of the active rest list (entry k at position
// k) so that a level's loads are coalesced, and read for the impact rest list through `pos`
// (contact -> position). h: [q * cap + k], q = normal + friction, tangent1 + restitution, tangent2 +
// rolling resistance, plane + rolling radius, rolling impulse + (count | planar << 3); p: per point
// m [(m * 3 + q) * cap + k], q = anchor a + adjusted separation, anchor b + relative velocity,
// normal impulse + tangent impulse; peak: [m * cap + k]; ax: a planar contact's axes per point m
// (makePlanarAxes, computed here once as the reference computes them at prepare and rebase),
// [2m * cap + k] = normal arms A, B, normal mass, tangent arm A; [(2m + 1) * cap + k] = tangent arm B,
// tangent mass. Every slot of the four points exists for every contact, so the solves load them
// all at once, without waiting for the count. The contacts themselves get the impulses back before a
// refresh reads them and at the end of the step; the pack is built again after a refresh.
struct RestPack {
    float4* h;
    float4* p;
    float* peak;
    float4* ax;
    float4* pre;     // a planar contact's PlanarRowPre for the current pass: [q * cap + k], q = scale, bias,
                     // impulse scale, one float4 lane per point
    uint32_t* pos;
    uint32_t cap;
};

// The impact rest entries carry their contact's pack position in entry.w >> 4 (k_rest_impact_pos).
constexpr uint32_t kRestPosShift = 4;

__global__ void k_rest_pack(const uint4* entries, uint32_t n, const tfs_contact* contacts, const float4* mass,
                            RestPack rp) {
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const uint4 e = entries[k];
        const uint32_t ci = e.x;
        const tfs_contact& c = contacts[ci];
        const size_t cap = rp.cap;
        const uint32_t meta = (uint32_t)c.count | (c.planar ? 8u : 0u);
        rp.h[k] = make_float4(c.normal[0], c.normal[1], c.normal[2], c.friction);
        rp.h[cap + k] = make_float4(c.tangent1[0], c.tangent1[1], c.tangent1[2], c.restitution);
        rp.h[2 * cap + k] = make_float4(c.tangent2[0], c.tangent2[1], c.tangent2[2], c.rolling_resistance);
        rp.h[3 * cap + k] = make_float4(c.plane[0], c.plane[1], c.plane[2], c.rolling_radius);
        rp.h[4 * cap + k] = make_float4(c.rolling_impulse[0], c.rolling_impulse[1], c.rolling_impulse[2], __uint_as_float(meta));
        TF_POINTS(m, c.count) {
            const tfs_point& q = c.points[m];
            rp.p[(size_t)(m * 3) * cap + k] = make_float4(q.anchor_a[0], q.anchor_a[1], q.anchor_a[2], q.adjusted_separation);
            rp.p[(size_t)(m * 3 + 1) * cap + k] = make_float4(q.anchor_b[0], q.anchor_b[1], q.anchor_b[2], q.relative_velocity);
            rp.p[(size_t)(m * 3 + 2) * cap + k] =
                make_float4(q.normal_impulse, q.tangent_impulse[0], q.tangent_impulse[1], q.tangent_impulse[2]);
            rp.peak[(size_t)m * cap + k] = q.max_normal_impulse;
        }
        if (c.planar) {
            const Mass ma = loadPlanarMass(mass, e.y), mb = loadPlanarMass(mass, e.z);
            TF_POINTS(m, c.count) {
                const PlanarAxes x = planarAxes(ma, mb, c, m);
                rp.ax[(size_t)(2 * m) * cap + k] = make_float4(x.normalA, x.normalB, x.normalMass, x.tangentA);
                rp.ax[(size_t)(2 * m + 1) * cap + k] = make_float4(x.tangentB, x.tangentMass, 0.0f, 0.0f);
            }
        }
        rp.pos[ci] = k;
    }
}

__global__ void k_rest_impact_pos(uint4* entries, uint32_t n, const uint32_t* pos) {
    for (uint32_t k = gid(); k < n; k += gstride()) {
        uint4 e = entries[k];
        e.w = (e.w & ((1u << kRestPosShift) - 1u)) | (pos[e.x] << kRestPosShift);
        entries[k] = e;
    }
}

// What the rest solves change goes back to the contacts: point impulses and the rolling impulse.
__global__ void k_rest_unpack(const uint4* entries, uint32_t n, RestPack rp, tfs_contact* contacts) {
    for (uint32_t k = gid(); k < n; k += gstride()) {
        tfs_contact& c = contacts[entries[k].x];
        const size_t cap = rp.cap;
        const float4 ri = rp.h[4 * cap + k];
        c.rolling_impulse[0] = ri.x;
        c.rolling_impulse[1] = ri.y;
        c.rolling_impulse[2] = ri.z;
        TF_POINTS(m, c.count) {
            const float4 im = rp.p[(size_t)(m * 3 + 2) * cap + k];
            c.points[m].normal_impulse = im.x;
            c.points[m].tangent_impulse[0] = im.y;
            c.points[m].tangent_impulse[1] = im.z;
            c.points[m].tangent_impulse[2] = im.w;
            c.points[m].max_normal_impulse = rp.peak[(size_t)m * cap + k];
        }
    }
}

// A rest contact from the pack, in the fields the rest solves read (the others are left unset), and
// its planar axes. The slots of all four points are loaded whatever the count, so that no load waits
// for another; the solves read only the first count of them.
__device__ __forceinline__ void loadRest(const RestPack& rp, uint32_t k, tfs_contact& c, PlanarAxes* ax) {
    const size_t cap = rp.cap;
    const float4 a = rp.h[k], b = rp.h[cap + k], d = rp.h[2 * cap + k], e = rp.h[3 * cap + k], f = rp.h[4 * cap + k];
    float4 pa[4], pb[4], im[4], x0[4], x1[4];
    float peak[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        pa[m] = rp.p[(size_t)(m * 3) * cap + k];
        pb[m] = rp.p[(size_t)(m * 3 + 1) * cap + k];
        im[m] = rp.p[(size_t)(m * 3 + 2) * cap + k];
        peak[m] = rp.peak[(size_t)m * cap + k];
        x0[m] = rp.ax[(size_t)(2 * m) * cap + k];
        x1[m] = rp.ax[(size_t)(2 * m + 1) * cap + k];
    }
    const uint32_t meta = __float_as_uint(f.w);
    c.count = (int32_t)(meta & 7u);
    c.planar = (meta & 8u) ? 1 : 0;
    c.laned = 0;
    c.normal[0] = a.x; c.normal[1] = a.y; c.normal[2] = a.z; c.friction = a.w;
    c.tangent1[0] = b.x; c.tangent1[1] = b.y; c.tangent1[2] = b.z; c.restitution = b.w;
    c.tangent2[0] = d.x; c.tangent2[1] = d.y; c.tangent2[2] = d.z; c.rolling_resistance = d.w;
    c.plane[0] = e.x; c.plane[1] = e.y; c.plane[2] = e.z; c.rolling_radius = e.w;
    c.rolling_impulse[0] = f.x; c.rolling_impulse[1] = f.y; c.rolling_impulse[2] = f.z;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        tfs_point& q = c.points[m];
        q.anchor_a[0] = pa[m].x; q.anchor_a[1] = pa[m].y; q.anchor_a[2] = pa[m].z; q.adjusted_separation = pa[m].w;
        q.anchor_b[0] = pb[m].x; q.anchor_b[1] = pb[m].y; q.anchor_b[2] = pb[m].z; q.relative_velocity = pb[m].w;
        q.normal_impulse = im[m].x; q.tangent_impulse[0] = im[m].y; q.tangent_impulse[1] = im[m].z; q.tangent_impulse[2] = im[m].w;
        q.max_normal_impulse = peak[m];
        ax[m] = PlanarAxes{ x0[m].x, x0[m].y, x0[m].z, x0[m].w, x1[m].x, x1[m].y };
    }
}

__device__ __forceinline__ void storeRest(const tfs_contact& c, const RestPack& rp, uint32_t k) {
    const size_t cap = rp.cap;
    const float4 f = rp.h[4 * cap + k];
    rp.h[4 * cap + k] = make_float4(c.rolling_impulse[0], c.rolling_impulse[1], c.rolling_impulse[2], f.w);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        if (m < c.count) {
            const tfs_point& q = c.points[m];
            rp.p[(size_t)(m * 3 + 2) * cap + k] =
                make_float4(q.normal_impulse, q.tangent_impulse[0], q.tangent_impulse[1], q.tangent_impulse[2]);
            rp.peak[(size_t)m * cap + k] = q.max_normal_impulse;
        }
    }
}

// The warm start of the rest contacts, one body at a time: a warm-start push does not depend on the
// velocities, so each body applies its rest contacts' effects (warmRestSide) in list order by
// itself, after its coloured rows (warmStart after warmStartBatches, src/world.cpp:1672-1673). A
// body's rest contacts in list order are its entries in level order (contacts sharing a body that
// they write follow each other's levels). Entries: pack position << 1 | side, for the sides each
// contact writes.
__global__ void k_wrest_count(const uint4* entries, uint32_t n, uint32_t* deg) {
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const uint4 e = entries[k];
        if (e.w & 1u) { atomicAdd(&deg[e.y], 1u); }
        if (e.w & 2u) { atomicAdd(&deg[e.z], 1u); }
    }
}

__global__ void k_wrest_fill(const uint4* entries, uint32_t n, const uint32_t* start, uint32_t* cursor, uint32_t* ep) {
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const uint4 e = entries[k];
        if (e.w & 1u) { ep[start[e.y] + atomicAdd(&cursor[e.y], 1u)] = k << 1; }
        if (e.w & 2u) { ep[start[e.z] + atomicAdd(&cursor[e.z], 1u)] = (k << 1) | 1u; }
    }
}

__device__ __forceinline__ void warmBodyRest(uint32_t d, const uint32_t* __restrict__ list, const uint4* __restrict__ entries,
                                             const RestPack& rp, const Mass& mx, Motion& m) {
    for (uint32_t j = 0; j < d; ++j) {
        const uint32_t en = list[j], k = en >> 1;
        const bool sideB = en & 1u;
        const uint32_t flags = entries[k].w;
        tfs_contact c;
        PlanarAxes ax[4];
        loadRest(rp, k, c, ax);
        warmRestSide(c, sideB, !(flags & (sideB ? 8u : 4u)), mx, m, ax);
    }
}

// The start of a substep, one thread per body: integrateVelocities (src/world.cpp:1141-1151), then the
// warm-start pushes of the body's coloured rows in row order (warmStartBatches), then those of its rest
// contacts in list order (warmStart of the rest, src/solver.cpp:319-358; REST). A body's velocity after
// them depends only on its own operations, in this order, so the three stages run as one.
template <bool REST>
__global__ void __launch_bounds__(kThreads)
    k_substep_start(tfs_body* bodies, uint32_t nb, float4* __restrict__ motion, Ctx ctx, const uint32_t* __restrict__ deg,
                    const uint32_t* __restrict__ start, const uint32_t* __restrict__ ep, Rows rows,
                    const uint32_t* __restrict__ restDeg, const uint32_t* __restrict__ restStart,
                    const uint32_t* __restrict__ restEp, const uint4* __restrict__ entries, RestPack rp,
                    const float4* __restrict__ mass, int parts) {
    // parts: 1 integrateVelocities, 2 the warm start, 3 both (with joints, their warm start goes between)
    for (uint32_t x = gid(); x < nb; x += gstride()) {
        tfs_body& body = bodies[x];
        const bool integrate = (parts & 1) && body.alive && body.type == TFS_DYNAMIC && body.awake;
        const uint32_t d = (parts & 2) ? deg[x] : 0u;
        const uint32_t dr = REST && (parts & 2) ? restDeg[x] : 0u;
        if (!integrate && d == 0 && dr == 0) { continue; }
        Motion m = loadMotion(motion, x);
        if (integrate) {
            float kept = body.kept;
            integrateVelocity(body, ctx, m.v, m.w, kept);
            body.kept = kept;
        }
        if (d > 0) { warmBodyRows(d, ep + start[x], rows, m); }
        if (REST && dr > 0) { warmBodyRest(dr, restEp + restStart[x], entries, rp, loadMass(mass, x), m); }
        storeVelocity(motion, x, m);
    }
}

// ---- the fused substep -------------------------------------------------------------------------------
//
// A whole substep without a refresh in one persistent launch (src/world.cpp:1671-1680): the substep
// start over the bodies (integrateVelocities and the warm start of every coloured row, as
// k_substep_start), the biased pass, the positions (as k_integrate_positions), the relax pass, with a
// grid barrier between the phases. Data that a phase writes and a later phase reads from another block
// (velocities, dp and dq, impulses) goes through coherent accesses, as within a pass; each body stays
// with one thread in both body phases. It needs every active contact in a colour (no rest list), and
// all its blocks resident at once, as the pass launches do. The operations and their order are those
// of the separate launches, so the results are the same bits.
template <bool MADE>
__global__ void __launch_bounds__(kPassThreads)
    k_substep_fused(tfs_body* bodies, uint32_t nb, Rows rows, const uint4* impactHdr, PassPlan bias, PassPlan relax,
                    float4* motion, const float4* mass, Ctx ctx, const uint32_t* __restrict__ deg,
                    const uint32_t* __restrict__ start, const uint32_t* __restrict__ ep, BarrierLines* bar,
                    BarrierBase base, DevCounts* cnt) {
    using X = Access<true>;
    BarrierPhase phase(base);
    for (uint32_t x = gid(); x < nb; x += gstride()) {
        tfs_body& body = bodies[x];
        const bool integrate = body.alive && body.type == TFS_DYNAMIC && body.awake;
        const uint32_t d = deg[x];
        if (!integrate && d == 0) { continue; }
        Motion m = loadMotion(motion, x);
        if (integrate) {
            float kept = body.kept;
            integrateVelocity(body, ctx, m.v, m.w, kept);
            body.kept = kept;
        }
        if (d > 0) { warmBodyRows(d, ep + start[x], rows, m); }
        X::st(&motion[4 * x], make_float4(m.v.x, m.v.y, m.v.z, 0.0f));
        X::st(&motion[4 * x + 1], make_float4(m.w.x, m.w.y, m.w.z, 0.0f));
    }
    gridSync(bar, phase, cnt);
    passRun<kBias, MADE, false, true>(rows, impactHdr, bias, motion, mass, ctx, bar, phase, cnt, Tape{});
    gridSync(bar, phase, cnt);
    for (uint32_t x = gid(); x < nb; x += gstride()) {
        tfs_body& b = bodies[x];
        if (!b.alive) { continue; }
        const float4 v = X::ld(&motion[4 * x]), w = X::ld(&motion[4 * x + 1]);
        const float4 dp = motion[4 * x + 2], dq = motion[4 * x + 3];
        Motion m{ vec(v.x, v.y, v.z), vec(w.x, w.y, w.z), vec(dp.x, dp.y, dp.z), quat(dq.x, dq.y, dq.z, dq.w) };
        float kept = b.kept;
        limitSpeed(b, m.v, m.w, kept, ctx.maxSpeed, ctx.maxSpin);
        b.kept = kept;
        if (b.type != TFS_STATIC && !(b.type == TFS_DYNAMIC && !b.awake)) { integratePosition(b, m, ctx.h); }
        X::st(&motion[4 * x], make_float4(m.v.x, m.v.y, m.v.z, 0.0f));
        X::st(&motion[4 * x + 1], make_float4(m.w.x, m.w.y, m.w.z, 0.0f));
        X::st(&motion[4 * x + 2], make_float4(m.dp.x, m.dp.y, m.dp.z, 0.0f));
        X::st(&motion[4 * x + 3], make_float4(m.dq.x, m.dq.y, m.dq.z, m.dq.w));
    }
    gridSync(bar, phase, cnt);
    passRun<kRelax, MADE, false, true>(rows, impactHdr, relax, motion, mass, ctx, bar, phase, cnt, Tape{});
}

// The rest list level after level. ONE_BLOCK: one work-group walks the levels with __syncthreads
// between them (narrow levels: the block barrier is cheap and one CU's accesses need no coherence
// protocol); otherwise one launch over many blocks with the grid barrier, velocities crossing blocks through
// coherent accesses as in k_pass. Each thread loads its next entry at the start of a level. A contact
// is touched by one thread only in a launch, so plain accesses serve for it.
constexpr int kRestBlockThreads = 512;

template <int MODE, bool ONE_BLOCK>
__global__ void __launch_bounds__(ONE_BLOCK ? kRestBlockThreads : kPassThreads)
    k_rest_levels(const uint4* entries, const uint32_t* levelStart, uint32_t levels, RestPack rp, bool viaPos,
                  float4* motion, const float4* mass, Ctx ctx, BarrierLines* bar, BarrierBase base, DevCounts* cnt) {
    using X = Access<!ONE_BLOCK>;
    BarrierPhase phase(base);
    const uint32_t g = ONE_BLOCK ? threadIdx.x : passSlot();
    const uint32_t stride = ONE_BLOCK ? blockDim.x : gstride();
    uint32_t lo = levelStart[0], hi = levels > 0 ? levelStart[1] : 0u;
    uint4 next = lo + g < hi ? entries[lo + g] : make_uint4(0u, 0u, 0u, 0u);
    for (uint32_t lv = 0; lv < levels; ++lv) {
        const uint32_t first = lo + g, end = hi;
        const uint4 firstEntry = next;
        if (lv + 1 < levels) {
            lo = end;
            hi = levelStart[lv + 2];
            if (lo + g < hi) { next = entries[lo + g]; }
        }
        for (uint32_t i = first; i < end; i += stride) {
            const uint4 e = i == first ? firstEntry : entries[i];
            const uint32_t k = viaPos ? e.w >> kRestPosShift : i;
            tfs_contact c;
            PlanarAxes ax[4];
            loadRest(rp, k, c, ax);
            const uint32_t ia = e.y, ib = e.z;
            const V3 none = vec(0.0f, 0.0f, 0.0f);
            const BodyView va{ (uint8_t)((e.w & 4u) ? TFS_STATIC : TFS_DYNAMIC), 1, 0, none };
            const BodyView vb{ (uint8_t)((e.w & 8u) ? TFS_STATIC : TFS_DYNAMIC), 1, 0, none };
            const float4 v0 = X::ld(&motion[4 * ia]), w0 = X::ld(&motion[4 * ia + 1]);
            const float4 v1 = X::ld(&motion[4 * ib]), w1 = X::ld(&motion[4 * ib + 1]);
            const float4 p0 = motion[4 * ia + 2], q0 = motion[4 * ia + 3], p1 = motion[4 * ib + 2], q1 = motion[4 * ib + 3];
            Motion a{ vec(v0.x, v0.y, v0.z), vec(w0.x, w0.y, w0.z), vec(p0.x, p0.y, p0.z), quat(q0.x, q0.y, q0.z, q0.w) };
            Motion b{ vec(v1.x, v1.y, v1.z), vec(w1.x, w1.y, w1.z), vec(p1.x, p1.y, p1.z), quat(q1.x, q1.y, q1.z, q1.w) };
            // a planar contact reads only the inverse mass and the planar inverse mass and inertia
            const Mass ma = c.planar ? loadPlanarMass(mass, ia) : loadMass(mass, ia);
            const Mass mb = c.planar ? loadPlanarMass(mass, ib) : loadMass(mass, ib);
            if (MODE == kRestitute) {
                restituteRest(c, va, vb, ma, mb, a, b, ctx, ax);
            } else if (c.planar) {
                const size_t cap = rp.cap;
                const float4 q0 = rp.pre[k], q1 = rp.pre[cap + k], q2 = rp.pre[2 * cap + k];
                const PlanarRowPre pre{ { q0.x, q0.y, q0.z, q0.w }, { q1.x, q1.y, q1.z, q1.w }, { q2.x, q2.y, q2.z, q2.w } };
                solvePlanar<MODE == kBias, true>(c, va, vb, ma, mb, a, b, ctx, ax, &pre);
            } else {
                solveRest<MODE == kBias>(c, va, vb, ma, mb, a, b, ctx, ax);
            }
            storeRest(c, rp, k);
            if (e.w & 1u) {
                X::st(&motion[4 * ia], make_float4(a.v.x, a.v.y, a.v.z, 0.0f));
                X::st(&motion[4 * ia + 1], make_float4(a.w.x, a.w.y, a.w.z, 0.0f));
            }
            if (e.w & 2u) {
                X::st(&motion[4 * ib], make_float4(b.v.x, b.v.y, b.v.z, 0.0f));
                X::st(&motion[4 * ib + 1], make_float4(b.w.x, b.w.y, b.w.z, 0.0f));
            }
        }
        if (lv + 1 < levels) {
            if (ONE_BLOCK) {
                __syncthreads();
            } else {
                gridSync(bar, phase, cnt);
            }
        }
    }
}

// The narrow levels of the active rest list in one work-group, with each level's contact data loaded
// while the level before it solves. A level's planar contacts read only what the pass does not
// change (directions, friction, rolling terms, their points' impulses and planar axes, the biases and
// factors of k_rest_pre, the bodies' planar masses) besides their bodies' velocities, so a thread
// loads its next level's contact into registers (a `WalkRow`) before solving the current one, and a
// level's dependent chain is its velocities' round trip, the solve and the stores. Contacts that are
// not planar, and the second and later contacts of a thread in a level wider than the work-group, take
// the general path (loadRest, solveRest) as k_rest_levels does. Pack positions are list positions.
constexpr uint32_t kWalkMaxLevels = 6144;  // level starts kept in LDS (24 KiB)
constexpr int kWalkThreads = 256;          // one wave per SIMD: room for two levels' rows in registers

struct WalkRow {
    uint4 e;
    float4 n, t1, plane, roll;
    float rollRes;
    float4 imp[4];
    float peak[4];
    float4 ax0[4];
    float2 ax1[4];
    float4 pre[3];
    float4 massA, massB;   // the mass records' planar parts (invMass, pim in .z, .w; pii in the next .x)
    float piiA, piiB;
};

__device__ __forceinline__ void loadWalkRow(const uint4* entries, const RestPack& rp, const float4* mass, uint32_t k,
                                            WalkRow& w) {
    const size_t cap = rp.cap;
    w.e = entries[k];
    w.n = rp.h[k];
    w.t1 = rp.h[cap + k];
    w.rollRes = rp.h[2 * cap + k].w;
    w.plane = rp.h[3 * cap + k];
    w.roll = rp.h[4 * cap + k];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        w.imp[m] = rp.p[(size_t)(m * 3 + 2) * cap + k];
        w.peak[m] = rp.peak[(size_t)m * cap + k];
        w.ax0[m] = rp.ax[(size_t)(2 * m) * cap + k];
        const float4 x1 = rp.ax[(size_t)(2 * m + 1) * cap + k];
        w.ax1[m] = make_float2(x1.x, x1.y);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) { w.pre[q] = rp.pre[(size_t)q * cap + k]; }
    w.massA = mass[6 * w.e.y + 4];
    w.massB = mass[6 * w.e.z + 4];
    w.piiA = mass[6 * w.e.y + 5].x;
    w.piiB = mass[6 * w.e.z + 5].x;
}

// One rest contact by the general path (k_rest_levels' body), for what the walk does not prefetch.
template <int MODE>
__device__ __forceinline__ void restGeneral(const uint4 e, uint32_t k, const RestPack& rp, float4* motion,
                                            const float4* mass, const Ctx& ctx) {
    tfs_contact c;
    PlanarAxes ax[4];
    loadRest(rp, k, c, ax);
    const uint32_t ia = e.y, ib = e.z;
    const V3 none = vec(0.0f, 0.0f, 0.0f);
    const BodyView va{ (uint8_t)((e.w & 4u) ? TFS_STATIC : TFS_DYNAMIC), 1, 0, none };
    const BodyView vb{ (uint8_t)((e.w & 8u) ? TFS_STATIC : TFS_DYNAMIC), 1, 0, none };
    Motion a = loadMotion(motion, ia), b = loadMotion(motion, ib);
    const Mass ma = c.planar ? loadPlanarMass(mass, ia) : loadMass(mass, ia);
    const Mass mb = c.planar ? loadPlanarMass(mass, ib) : loadMass(mass, ib);
    if (c.planar) {
        const size_t cap = rp.cap;
        const float4 q0 = rp.pre[k], q1 = rp.pre[cap + k], q2 = rp.pre[2 * cap + k];
        const PlanarRowPre pre{ { q0.x, q0.y, q0.z, q0.w }, { q1.x, q1.y, q1.z, q1.w }, { q2.x, q2.y, q2.z, q2.w } };
        solvePlanar<MODE == kBias, true>(c, va, vb, ma, mb, a, b, ctx, ax, &pre);
    } else {
        solveRest<MODE == kBias>(c, va, vb, ma, mb, a, b, ctx, ax);
    }
    storeRest(c, rp, k);
    if (e.w & 1u) {
        motion[4 * ia] = make_float4(a.v.x, a.v.y, a.v.z, 0.0f);
        motion[4 * ia + 1] = make_float4(a.w.x, a.w.y, a.w.z, 0.0f);
    }
    if (e.w & 2u) {
        motion[4 * ib] = make_float4(b.v.x, b.v.y, b.v.z, 0.0f);
        motion[4 * ib + 1] = make_float4(b.w.x, b.w.y, b.w.z, 0.0f);
    }
}

// One level of the walk for thread g: its contact `cur` (from the level before), the next level's row
// loaded into `next` meanwhile, then the level's other contacts of this thread by the general path.
template <int MODE>
__device__ __forceinline__ void walkLevel(uint32_t lv, uint32_t levels, const uint32_t* starts, uint32_t g,
                                          const uint4* entries, const RestPack& rp, float4* motion, const float4* mass,
                                          const Ctx& ctx, const WalkRow& cur, bool have, WalkRow& next, bool& haveNext) {
    const uint32_t lo = starts[lv], hi = starts[lv + 1];
    const bool planar = have && (__float_as_uint(cur.roll.w) & 8u);
    // this level's velocities first, then the next level's row: loads complete in the order they
    // are issued, so the solve waits only for its own
    const uint32_t ia = cur.e.y, ib = cur.e.z;
    float4 v0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), w0 = v0, v1 = v0, w1 = v0;
    if (planar) {
        v0 = motion[4 * ia];
        w0 = motion[4 * ia + 1];
        v1 = motion[4 * ib];
        w1 = motion[4 * ib + 1];
    }
    haveNext = false;
    if (lv + 1 < levels) {
        haveNext = hi + g < starts[lv + 2];
        if (haveNext) { loadWalkRow(entries, rp, mass, hi + g, next); }
    }
    if (have) {
        const uint32_t k = lo + g;
        if (planar) {
            const uint32_t meta = __float_as_uint(cur.roll.w);
            tfs_contact c;
            c.count = (int32_t)(meta & 7u);
            c.planar = 1;
            c.normal[0] = cur.n.x; c.normal[1] = cur.n.y; c.normal[2] = cur.n.z; c.friction = cur.n.w;
            c.tangent1[0] = cur.t1.x; c.tangent1[1] = cur.t1.y; c.tangent1[2] = cur.t1.z;
            c.rolling_resistance = cur.rollRes;
            c.plane[0] = cur.plane.x; c.plane[1] = cur.plane.y; c.plane[2] = cur.plane.z; c.rolling_radius = cur.plane.w;
            c.rolling_impulse[0] = cur.roll.x; c.rolling_impulse[1] = cur.roll.y; c.rolling_impulse[2] = cur.roll.z;
            PlanarAxes ax[4];
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                tfs_point& q = c.points[m];
                q.normal_impulse = cur.imp[m].x;
                q.tangent_impulse[0] = cur.imp[m].y; q.tangent_impulse[1] = cur.imp[m].z; q.tangent_impulse[2] = cur.imp[m].w;
                q.max_normal_impulse = cur.peak[m];
                ax[m] = PlanarAxes{ cur.ax0[m].x, cur.ax0[m].y, cur.ax0[m].z, cur.ax0[m].w, cur.ax1[m].x, cur.ax1[m].y };
            }
            const PlanarRowPre pre{ { cur.pre[0].x, cur.pre[0].y, cur.pre[0].z, cur.pre[0].w },
                                    { cur.pre[1].x, cur.pre[1].y, cur.pre[1].z, cur.pre[1].w },
                                    { cur.pre[2].x, cur.pre[2].y, cur.pre[2].z, cur.pre[2].w } };
            Mass ma{}, mb{};
            ma.invMass = cur.massA.z; ma.pim = cur.massA.w; ma.pii = cur.piiA;
            mb.invMass = cur.massB.z; mb.pim = cur.massB.w; mb.pii = cur.piiB;
            const V3 none = vec(0.0f, 0.0f, 0.0f);
            const BodyView va{ (uint8_t)((cur.e.w & 4u) ? TFS_STATIC : TFS_DYNAMIC), 1, 0, none };
            const BodyView vb{ (uint8_t)((cur.e.w & 8u) ? TFS_STATIC : TFS_DYNAMIC), 1, 0, none };
            const Q4 unit = quat(0.0f, 0.0f, 0.0f, 1.0f);
            Motion a{ vec(v0.x, v0.y, v0.z), vec(w0.x, w0.y, w0.z), none, unit };
            Motion b{ vec(v1.x, v1.y, v1.z), vec(w1.x, w1.y, w1.z), none, unit };
            solvePlanar<MODE == kBias, true>(c, va, vb, ma, mb, a, b, ctx, ax, &pre);
            storeRest(c, rp, k);
            if (cur.e.w & 1u) {
                motion[4 * ia] = make_float4(a.v.x, a.v.y, a.v.z, 0.0f);
                motion[4 * ia + 1] = make_float4(a.w.x, a.w.y, a.w.z, 0.0f);
            }
            if (cur.e.w & 2u) {
                motion[4 * ib] = make_float4(b.v.x, b.v.y, b.v.z, 0.0f);
                motion[4 * ib + 1] = make_float4(b.w.x, b.w.y, b.w.z, 0.0f);
            }
        } else {
            restGeneral<MODE>(cur.e, k, rp, motion, mass, ctx);
        }
    }
    for (uint32_t i = lo + g + blockDim.x; i < hi; i += blockDim.x) { restGeneral<MODE>(entries[i], i, rp, motion, mass, ctx); }
    __syncthreads();
}

// The levels two at a time, the rows of alternate levels in two register sets (no copy between levels).
template <int MODE>
__global__ void __launch_bounds__(kWalkThreads)
    k_rest_walk(const uint4* entries, const uint32_t* levelStart, uint32_t levels, RestPack rp, float4* motion,
                const float4* mass, Ctx ctx) {
    __shared__ uint32_t starts[kWalkMaxLevels + 1];
    for (uint32_t i = threadIdx.x; i <= levels; i += blockDim.x) { starts[i] = levelStart[i]; }
    __syncthreads();
    const uint32_t g = threadIdx.x;
    WalkRow rowA, rowB;
    bool haveA = starts[0] + g < starts[1], haveB = false;
    if (haveA) { loadWalkRow(entries, rp, mass, starts[0] + g, rowA); }
    for (uint32_t lv = 0; lv < levels; lv += 2) {
        walkLevel<MODE>(lv, levels, starts, g, entries, rp, motion, mass, ctx, rowA, haveA, rowB, haveB);
        if (lv + 1 < levels) {
            walkLevel<MODE>(lv + 1, levels, starts, g, entries, rp, motion, mass, ctx, rowB, haveB, rowA, haveA);
        }
    }
}

// planarPre of every planar contact of the active rest list before a velocity pass walks its levels:
// the pass does not change dp and dq, so each point's bias and impulse factors leave the levels'
// dependent chain; the impact rest list, in the same pass, reads the same pack positions.
template <bool BIAS>
__global__ void __launch_bounds__(kThreads) k_rest_pre(const uint4* __restrict__ entries, uint32_t n, RestPack rp,
                                                       const float4* __restrict__ motion, Ctx ctx) {
    const size_t cap = rp.cap;
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const uint32_t meta = __float_as_uint(rp.h[4 * cap + k].w);
        if (!(meta & 8u)) { continue; }
        const uint4 e = entries[k];
        tfs_contact c;
        PlanarAxes ax[4];
        loadRest(rp, k, c, ax);
        const Motion a = loadMotion(motion, e.y), b = loadMotion(motion, e.z);
        PlanarRowPre q;
        planarPre(c, !(e.w & 4u), !(e.w & 8u), a, b, ctx, ax, BIAS, q);
        rp.pre[k] = make_float4(q.scale[0], q.scale[1], q.scale[2], q.scale[3]);
        rp.pre[cap + k] = make_float4(q.bias[0], q.bias[1], q.bias[2], q.bias[3]);
        rp.pre[2 * cap + k] = make_float4(q.impulseScale[0], q.impulseScale[1], q.impulseScale[2], q.impulseScale[3]);
    }
}

// integratePositions: limitSpeeds, then each moving body's pose (src/world.cpp:1161-1172).
__global__ void k_integrate_positions(tfs_body* bodies, uint32_t n, float4* motion, Ctx ctx) {
    for (uint32_t i = gid(); i < n; i += gstride()) {
        tfs_body& b = bodies[i];
        if (!b.alive) { continue; }
        Motion m = loadMotion(motion, i);
        float kept = b.kept;
        limitSpeed(b, m.v, m.w, kept, ctx.maxSpeed, ctx.maxSpin);
        b.kept = kept;
        if (b.type != TFS_STATIC && !(b.type == TFS_DYNAMIC && !b.awake)) { integratePosition(b, m, ctx.h); }
        storeMotion(motion, i, m);
    }
}

__global__ void k_limit_speeds(tfs_body* bodies, uint32_t n, float4* motion, Ctx ctx) {
    for (uint32_t i = gid(); i < n; i += gstride()) {
        tfs_body& b = bodies[i];
        if (!b.alive) { continue; }
        Motion m = loadMotion(motion, i);
        float kept = b.kept;
        limitSpeed(b, m.v, m.w, kept, ctx.maxSpeed, ctx.maxSpin);
        b.kept = kept;
        storeVelocity(motion, i, m);
    }
}

__global__ void k_writeback(tfs_body* bodies, uint32_t n, const float4* motion) {
    for (uint32_t i = gid(); i < n; i += gstride()) {
        const Motion m = loadMotion(motion, i);
        tfs_body& b = bodies[i];
        st3(b.v, m.v);
        st3(b.w, m.w);
        st3(b.dp, m.dp);
        st4(b.dq, m.dq);
    }
}

// ---- refresh: src/world.cpp:1660-1668, 1057-1139 --------------------------------------------------

// Before the narrowphase of a refresh, over the active contacts: the coloured rows' impulses go back
// into their contacts (the refresh carries them), and every contact that is not collided again is
// rebased onto the bodies' current poses (src/world.cpp:1061-1067). `again` marks the contacts
// collided again: the fast flags of the step (src/world.cpp:1659).
// With several worlds resident, only the worlds that have contacts to collide again take part in a
// refresh (the reference refreshes a world only then): `refreshing` flags them by world, and
// `bodyWorld` maps a body to its world; both are null for a single world.
__device__ __forceinline__ bool takesPart(const uint8_t* refreshing, const uint32_t* bodyWorld, uint32_t body) {
    return !refreshing || refreshing[bodyWorld[body]];
}

__global__ void k_refresh_worlds(const uint32_t* again, uint32_t n, const tfs_contact* contacts,
                                 const uint32_t* bodyWorld, uint8_t* refreshing) {
    for (uint32_t k = gid(); k < n; k += gstride()) { refreshing[bodyWorld[contacts[again[k]].body_a]] = 1; }
}

__global__ void k_refresh_pre(const uint32_t* active, const DevCounts* cnt, const uint32_t* againFlag,
                              tfs_contact* contacts, Rows rows, const uint32_t* rowOf, const uint8_t* colors,
                              const float4* motion, const uint8_t* refreshing, const uint32_t* bodyWorld) {
    const uint32_t n = cnt->nActive;
    for (uint32_t k = gid(); k < n; k += gstride()) {
        const uint32_t ci = active[k];
        tfs_contact& c = contacts[ci];
        // Every coloured row's impulses go back to its contact, in every world: all rows are packed
        // again after the refresh.
        if (colors[ci] < TFS_COLORS) {
            const uint32_t r = rowOf[ci];
            for (int i = 0; i < c.count; ++i) {
                const uint32_t j = (uint32_t)i * rows.cap + r;
                const float4 im = rows.imp[j];
                c.points[i].normal_impulse = im.x;
                c.points[i].tangent_impulse[0] = im.y;
                c.points[i].tangent_impulse[1] = im.z;
                c.points[i].tangent_impulse[2] = im.w;
                c.points[i].max_normal_impulse = rows.peak[j];
            }
        }
        if (!againFlag[ci] && takesPart(refreshing, bodyWorld, c.body_a)) {
            rebaseContact(loadMotion(motion, c.body_a), loadMotion(motion, c.body_b), c);
        }
    }
}

// Then over the bodies: the inverse mass of awake dynamic bodies (src/world.cpp:1662-1664), and every
// body's dp, dq and base orientation reset (src/world.cpp:1068-1072).
__global__ void k_refresh_bodies(tfs_body* bodies, uint32_t n, float4* motion, float4* mass,
                                 const uint8_t* refreshing, const uint32_t* bodyWorld) {
    for (uint32_t i = gid(); i < n; i += gstride()) {
        if (!takesPart(refreshing, bodyWorld, i)) { continue; }
        tfs_body& b = bodies[i];
        if (b.alive && b.type == TFS_DYNAMIC && b.awake) {
            updateInverseMass(b);
            storeMass(mass, i, b);
        }
        for (int k = 0; k < 4; ++k) { b.q_base[k] = b.q[k]; }
        motion[4 * i + 2] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        motion[4 * i + 3] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    }
}

// The contacts to collide again and their two bodies (velocities and travel from the motion records),
// one wave per contact copying word by word, straight into compact host arrays in the list's order:
// hostC[k] and hostB[2k], hostB[2k + 1].
__global__ void k_gather_again(const uint32_t* again, uint32_t n, const tfs_contact* contacts, const tfs_body* bodies,
                               const float4* motion, uint32_t* hostAgain, tfs_contact* hostC, tfs_body* hostB) {
    static_assert(sizeof(tfs_contact) % 4 == 0 && sizeof(tfs_body) % 4 == 0, "word copies");
    constexpr uint32_t cw = sizeof(tfs_contact) / 4, bw = sizeof(tfs_body) / 4;
    constexpr uint32_t ov = offsetof(tfs_body, v) / 4, ow = offsetof(tfs_body, w) / 4;
    constexpr uint32_t odp = offsetof(tfs_body, dp) / 4, odq = offsetof(tfs_body, dq) / 4;
    const uint32_t lane = threadIdx.x % 64, waves = gstride() / 64;
    for (uint32_t k = gid() / 64; k < n; k += waves) {
        const uint32_t ci = again[k];
        if (lane == 0) { hostAgain[k] = ci; }
        const uint32_t* src = (const uint32_t*)(contacts + ci);
        uint32_t* dst = (uint32_t*)(hostC + k);
        for (uint32_t d = lane; d < cw; d += 64) { dst[d] = src[d]; }
        const uint32_t ids[2] = { contacts[ci].body_a, contacts[ci].body_b };
        for (int e = 0; e < 2; ++e) {
            const uint32_t i = ids[e];
            const uint32_t* bs = (const uint32_t*)(bodies + i);
            const uint32_t* ms = (const uint32_t*)(motion + 4 * (size_t)i);
            uint32_t* bd = (uint32_t*)(hostB + 2 * (size_t)k + e);
            for (uint32_t d = lane; d < bw; d += 64) {
                uint32_t x = bs[d];
                if (d >= ov && d < ov + 3) {
                    x = ms[d - ov];
                } else if (d >= ow && d < ow + 3) {
                    x = ms[4 + d - ow];
                } else if (d >= odp && d < odp + 3) {
                    x = ms[8 + d - odp];
                } else if (d >= odq && d < odq + 4) {
                    x = ms[12 + d - odq];
                }
                bd[d] = x;
            }
        }
    }
}

// The narrowphase half of a refresh on the device (src/world.cpp:1076-1122 for one shape per body): each
// listed pair collided again at its bodies' current poses and velocities over the time left in the step
// with the reference's narrowphase (collide3d.h: spheres, boxes, half-spaces), into fresh[k] as the
// refresh callback would write it. A pair collide3d.h does not collide (another kind of shape, or a pair
// the reference collides in 2D) raises error bit 16.
//
// A wave takes as long as the costliest path any of its lanes takes (two boxes: the closed-form
// queries, the gap test in double, clipping), so the pairs go in `order`: the list's positions with
// the pairs of two boxes first, then sphere-box, half-space-box and the rest (k_refresh_order), one
// wave to a work-group so the waves spread over the CUs. Without an order (a list longer than
// kOrderMax) the pairs go in list order. At one wave per SIMD the compiler keeps the clipping's and the
// gap test's small arrays (indexed at run time) in LDS rather than in scratch memory.
__global__ void __launch_bounds__(kCollideThreads) __attribute__((amdgpu_waves_per_eu(1, 1))) k_refresh_collide(const uint32_t* again, const uint32_t* order, uint32_t n,
                                                                     const tfs_contact* contacts, const tfs_body* bodies,
                                                                     const float4* motion, const tfs_shape* shapes,
                                                                     float remaining, float speculative, float linearSlop,
                                                                     tfs_contact* fresh, DevCounts* cnt) {
    for (uint32_t j = gid(); j < n; j += gstride()) {
        const uint32_t k = order ? order[j] : j;
        const tfs_contact& c = contacts[again[k]];
        const uint32_t ia = c.body_a, ib = c.body_b;
        const tfs_body& a = bodies[ia];
        const tfs_body& b = bodies[ib];
        const tfs_shape sa = shapes[ia], sb = shapes[ib];
        const float4 va = motion[4 * ia], wa = motion[4 * ia + 1], vb = motion[4 * ib], wb = motion[4 * ib + 1];
        const tfc::Solid solidA{ (int)sa.kind, ld3(sa.half), sa.radius }, solidB{ (int)sb.kind, ld3(sb.half), sb.radius };
        tfc::Manifold m;
        const int got = tfc::pairManifold(a, vec(va.x, va.y, va.z), vec(wa.x, wa.y, wa.z), solidA, b, vec(vb.x, vb.y, vb.z),
                                          vec(wb.x, wb.y, wb.z), solidB, remaining, speculative, linearSlop, m);
        tfs_contact& f = fresh[k];
        f.count = 0;
        if (got < 0) {
            atomicOr(&cnt->error, 16u);
        } else if (got > 0) {
            V3 pos[4];
            float sep[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                pos[i] = m.points[i].position;
                sep[i] = m.points[i].separation;
            }
            tfc::fillContact(a, b, m.normal, m.count, pos, sep, speculative, f);
        }
    }
}

// The cost class of a pair for the collision order: 0 two boxes, 1 a sphere and a box, 2 a half-space and
// a box, 3 anything else (two spheres, a half-space and a sphere, kinds the device does not collide).
__device__ __forceinline__ uint32_t collideClass(uint32_t ka, uint32_t kb) {
    const bool boxA = ka == (uint32_t)tfc::kBox, boxB = kb == (uint32_t)tfc::kBox;
    if (boxA && boxB) { return 0; }
    const bool sphere = ka == (uint32_t)tfc::kSphere || kb == (uint32_t)tfc::kSphere;
    const bool half = ka == (uint32_t)tfc::kHalfSpace || kb == (uint32_t)tfc::kHalfSpace;
    if ((boxA || boxB) && sphere) { return 1; }
    if ((boxA || boxB) && half) { return 2; }
    return 3;
}

// The collision order of the two fast lists (work-group 0: the fast list, 1: the fast solid list), by
// cost class, costliest first: a counting sort in LDS. Within a class the order is whatever the atomics
// give; each pair's manifold goes to its own list position, so the results do not depend on it.
// Lists longer than kOrderMax are left in list order (the host passes no order for them).
__global__ void __launch_bounds__(1024) k_refresh_order(const uint32_t* fast, const uint32_t* fastSolid, const DevCounts* cnt,
                                                        const tfs_contact* contacts, const tfs_shape* shapes,
                                                        uint32_t* fastOrder, uint32_t* fastSolidOrder) {
    constexpr uint32_t kClasses = 4;
    __shared__ uint32_t count[kClasses];
    const uint32_t* list = blockIdx.x ? fastSolid : fast;
    uint32_t* out = blockIdx.x ? fastSolidOrder : fastOrder;
    const uint32_t n = blockIdx.x ? cnt->nFastSolid : cnt->nFast;
    if (n > kOrderMax) { return; }
    if (threadIdx.x < kClasses) { count[threadIdx.x] = 0; }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        const tfs_contact& c = contacts[list[k]];
        atomicAdd(&count[collideClass(shapes[c.body_a].kind, shapes[c.body_b].kind)], 1u);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t at = 0;
        for (uint32_t q = 0; q < kClasses; ++q) {
            const uint32_t m = count[q];
            count[q] = at;
            at += m;
        }
    }
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        const tfs_contact& c = contacts[list[k]];
        out[atomicAdd(&count[collideClass(shapes[c.body_a].kind, shapes[c.body_b].kind)], 1u)] = k;
    }
}

// The solver side of each refreshed fast contact (src/world.cpp:1108-1138).
// A contact collided again is no longer reusable and its points' surface points are cleared
// (src/world.cpp:1110, 1117: the points start over), in its reuse data `ext`.
// Each refreshed contact keeps its row, whose points and axes are then packed again (as k_pack_again,
// by the same thread: the row reads only its own contact).
__device__ __forceinline__ void refreshOne(uint32_t k, const uint32_t* again, const tfs_contact* fresh, tfs_contact* contacts,
                                           const tfs_body* bodies, const float4* motion, const float4* mass, float matchSq,
                                           tfs_contact_ext* ext) {
        tfs_contact& c = contacts[again[k]];
        const tfs_contact& f = fresh[k];
        if (f.count <= 0) {
            c.count = 0;
            return;
        }
        {
            tfs_contact_ext& e = ext[again[k]];
            e.reusable = 0;
            const int m = f.count > 4 ? 4 : f.count;
            for (int i = 0; i < m; ++i) {
                for (int j = 0; j < 3; ++j) {
                    e.surface_a[i][j] = 0.0f;
                    e.surface_b[i][j] = 0.0f;
                }
            }
        }
        // The contact as it was (read in place: c is written once, at the end) and as it is refreshed.
        const tfs_contact& old = c;
        tfs_contact x = c;
        for (int j = 0; j < 3; ++j) {
            x.normal[j] = f.normal[j];
            x.tangent1[j] = f.tangent1[j];
            x.tangent2[j] = f.tangent2[j];
        }
        x.count = f.count > 4 ? 4 : f.count;
        TF_POINTS(i, x.count) {
            tfs_point& p = x.points[i];
            p = tfs_point{};
            for (int j = 0; j < 3; ++j) {
                p.local_a[j] = f.points[i].local_a[j];
                p.anchor_a[j] = f.points[i].anchor_a[j];
                p.anchor_b[j] = f.points[i].anchor_b[j];
            }
            p.separation = f.points[i].separation;
        }
        const tfs_body& ta = bodies[x.body_a];
        const tfs_body& tb = bodies[x.body_b];
        prepareContact(viewOf(ta), viewOf(tb), ta, tb, loadMotion(motion, x.body_a), loadMotion(motion, x.body_b),
                       loadMass(mass, x.body_a), loadMass(mass, x.body_b), x);
        TF_POINTS(i, x.count) {
            const int near = nearestPoint(old, ld3(x.points[i].local_a));
            if (near < 0) { continue; }
            x.points[i].relative_velocity = old.points[near].relative_velocity;
            x.points[i].max_normal_impulse = old.points[near].max_normal_impulse;
        }
        carryImpulses(old, x, matchSq);
        c = x;
}

__global__ void __launch_bounds__(kThreads) k_refresh_apply(const uint32_t* again, uint32_t n, const tfs_contact* fresh, tfs_contact* contacts,
                                const tfs_body* bodies, const float4* motion, const float4* mass, float matchSq,
                                tfs_contact_ext* ext, const uint8_t* colors, const uint8_t* moving, Rows rows,
                                uint32_t* rowOf) {
    for (uint32_t k = gid(); k < n; k += gstride()) {
        refreshOne(k, again, fresh, contacts, bodies, motion, mass, matchSq, ext);
        const uint32_t ci = again[k];
        if (colors[ci] < TFS_COLORS) { packRow(rowOf[ci], ci, contacts, moving, mass, rows, rowOf); }
    }
}

// ---- updateSleep: src/world.cpp:1174-1247 ---------------------------------------------------------

__global__ void k_sleep_time(tfs_body* bodies, uint32_t n, float dt, float threshold, uint32_t* parent, uint32_t* islandTime) {
    for (uint32_t i = gid(); i < n; i += gstride()) {
        parent[i] = i;
        islandTime[i] = 0x7f7fffffu;  // FLT_MAX
        tfs_body& b = bodies[i];
        if (!b.alive || b.type != TFS_DYNAMIC || !b.awake) { continue; }
        const float speed = maxf(length(ld3(b.v)), length(ld3(b.w)) * b.extent);
        if (!b.can_sleep || speed > threshold) {
            b.sleep_time = 0.0f;
        } else {
            b.sleep_time += dt;
        }
    }
}

__device__ __forceinline__ bool awakeDynamic(const tfs_body& b) { return b.type == TFS_DYNAMIC && b.awake; }

// Kinematic bodies that move keep their dynamic partners awake. Also writes the island graph's
// edges: edges[k] is the contact's two bodies when both are awake and dynamic, else (0, 0).
__global__ void k_sleep_kinematic(const tfs_contact* contacts, uint32_t nc, tfs_body* bodies, uint2* edges) {
    for (uint32_t k = gid(); k < nc; k += gstride()) {
        const tfs_contact& c = contacts[k];
        edges[k] = make_uint2(0u, 0u);
        if (c.count == 0) { continue; }
        tfs_body& a = bodies[c.body_a];
        tfs_body& b = bodies[c.body_b];
        const bool da = awakeDynamic(a), db = awakeDynamic(b);
        if (da && db) {
            edges[k] = make_uint2(c.body_a, c.body_b);
            continue;
        }
        if (da && b.type == TFS_KINEMATIC && movingAwake(b, ld3(b.v), ld3(b.w))) {
            a.sleep_time = 0.0f;
        } else if (db && a.type == TFS_KINEMATIC && movingAwake(a, ld3(a.v), ld3(a.w))) {
            b.sleep_time = 0.0f;
        }
    }
}

// Parents only ever move to smaller indices of the same island, so a stale parent read within a
// launch is still an ancestor: plain loads are enough, and the next round sees every update.
__device__ __forceinline__ uint32_t rootOf(const uint32_t* parent, uint32_t x) {
    uint32_t p = parent[x];
    while (p != x) {
        x = p;
        p = parent[x];
    }
    return x;
}

// Links the islands of two awake dynamic bodies in contact. Every link points from a larger index
// to a smaller one, so each island ends with its lowest body index as root, as the reference's
// union by the smaller root gives (src/world.cpp:1193-1194, include/twofold/world.h:464-465).
// A round given `prev` runs only if the round before it linked something (*prev != 0): rounds after
// one that links nothing change nothing, so they can be launched ahead of the check.
__global__ void k_sleep_hook(const uint2* edges, uint32_t nc, uint32_t* parent, uint32_t* changed, const uint32_t* prev) {
    if (prev && *prev == 0u) { return; }
    bool linked = false;
    for (uint32_t k = gid(); k < nc; k += gstride()) {
        const uint2 e = edges[k];
        if (e.x == e.y) { continue; }
        const uint32_t ra = rootOf(parent, e.x), rb = rootOf(parent, e.y);
        if (ra == rb) { continue; }
        atomicMin(&parent[ra > rb ? ra : rb], ra < rb ? ra : rb);
        linked = true;
    }
    if (__ballot(linked) != 0ull && (threadIdx.x % 64) == 0) { *changed = 1; }
}

__global__ void k_sleep_jump(uint32_t* parent, uint32_t n, const uint32_t* prev) {
    if (prev && *prev == 0u) { return; }
    for (uint32_t i = gid(); i < n; i += gstride()) { parent[i] = rootOf(parent, i); }
}

// The joints' part of the island pass (src/world.cpp:1210-1223), after the contacts' edges: a joint that
// is alive and not broken joins two awake dynamic bodies, or keeps a dynamic body awake beside a moving
// kinematic one. The world body is static.
__global__ void k_sleep_joints(const tfs_joint* joints, uint32_t nj, tfs_body* bodies, uint2* edges) {
    for (uint32_t k = gid(); k < nj; k += gstride()) {
        const tfs_joint& j = joints[k];
        edges[k] = make_uint2(0u, 0u);
        if (!j.alive || j.broken) { continue; }
        const bool worldA = j.body_a == TFS_WORLD_BODY;
        tfs_body& b = bodies[j.body_b];
        const bool da = !worldA && awakeDynamic(bodies[j.body_a]), db = awakeDynamic(b);
        if (da && db) {
            edges[k] = make_uint2(j.body_a, j.body_b);
            continue;
        }
        if (da && b.type == TFS_KINEMATIC && movingAwake(b, ld3(b.v), ld3(b.w))) {
            bodies[j.body_a].sleep_time = 0.0f;
        } else if (db && !worldA && bodies[j.body_a].type == TFS_KINEMATIC) {
            tfs_body& a = bodies[j.body_a];
            if (movingAwake(a, ld3(a.v), ld3(a.w))) { b.sleep_time = 0.0f; }
        }
    }
}

// ---- joints: src/world.cpp:1623-1632, 1647, 1670-1678, 1685-1691; src/joint.cpp ------------------------
//
// The reference solves its joints one at a time, in order, before the contacts of every pass, so each
// kernel here is one thread walking the active joints; a joint's two ends are loaded, solved and stored
// before the next joint reads them.

__device__ tfj::End loadEnd(uint32_t i, const tfs_body* bodies, const float4* motion, const float4* world) {
    tfj::End e;
    if (i == TFS_WORLD_BODY) {
        e.v = vec(world[0].x, world[0].y, world[0].z);
        e.w = vec(world[1].x, world[1].y, world[1].z);
        e.com = e.lc = e.p = vec(0.0f, 0.0f, 0.0f);
        e.q = quat(0.0f, 0.0f, 0.0f, 1.0f);
        e.iMW = e.iIW = tf::mzero();
        e.invMass = 0.0f;
        e.locks = 0;
        e.type = TFS_STATIC;
        e.awake = 1;
        return e;
    }
    const tfs_body& b = bodies[i];
    const float4 v = motion[4 * i], w = motion[4 * i + 1];
    e.v = vec(v.x, v.y, v.z);
    e.w = vec(w.x, w.y, w.z);
    e.com = ld3(b.com);
    e.lc = ld3(b.local_center);
    e.p = ld3(b.p);
    e.q = ld4(b.q);
    e.iMW = tf::ld9(b.inv_mass_world);
    e.iIW = tf::ld9(b.inv_inertia_world);
    e.invMass = b.inv_mass;
    e.locks = b.locks;
    e.type = b.type;
    e.awake = b.awake;
    return e;
}
__device__ void storeEndVelocity(uint32_t i, const tfj::End& e, float4* motion, float4* world) {
    float4* m = i == TFS_WORLD_BODY ? world : motion + 4 * (size_t)i;
    m[0] = make_float4(e.v.x, e.v.y, e.v.z, 0.0f);
    m[1] = make_float4(e.w.x, e.w.y, e.w.z, 0.0f);
}
__device__ void storeEndPose(uint32_t i, const tfj::End& e, tfs_body* bodies) {
    if (i == TFS_WORLD_BODY) { return; }
    tfs_body& b = bodies[i];
    st3(b.com, e.com);
    st4(b.q, e.q);
    st3(b.p, e.p);
}
__device__ __forceinline__ bool jointEndAwake(uint32_t i, const tfs_body* bodies) {
    return i != TFS_WORLD_BODY && bodies[i].type == TFS_DYNAMIC && bodies[i].awake;
}

// The step's active joints (alive, not broken, with an awake dynamic end), their force and torque cleared.
__global__ void k_joints_begin(tfs_joint* joints, uint32_t nj, const tfs_body* bodies, uint32_t* active, DevCounts* cnt) {
    if (gid() != 0) { return; }
    uint32_t n = 0;
    for (uint32_t k = 0; k < nj; ++k) {
        tfs_joint& j = joints[k];
        if (!j.alive || j.broken) { continue; }
        if (!jointEndAwake(j.body_a, bodies) && !jointEndAwake(j.body_b, bodies)) { continue; }
        active[n++] = k;
        st3(j.force, vec(0.0f, 0.0f, 0.0f));
        st3(j.torque, vec(0.0f, 0.0f, 0.0f));
    }
    cnt->nJointsActive = n;
    cnt->nJointsBegun = n;
}

// A substep's joint start, after integrateVelocities: scaleToKeptSpeed, then warmStartJoints.
__global__ void k_joints_warm(tfs_joint* joints, const uint32_t* active, const DevCounts* cnt, tfs_body* bodies,
                              float4* motion, float4* world) {
    if (gid() != 0) { return; }
    const uint32_t n = cnt->nJointsActive;
    auto kept = [&](uint32_t i) { return i == TFS_WORLD_BODY ? 1.0f : bodies[i].kept; };
    for (uint32_t k = 0; k < n; ++k) {
        tfs_joint& j = joints[active[k]];
        const float f = minf(kept(j.body_a), kept(j.body_b));
        if (f >= 1.0f) { continue; }
        tfj::scaleImpulses(j, f);
    }
    for (uint32_t k = 0; k < n; ++k) {
        const tfs_joint& j = joints[active[k]];
        if (j.body_a != TFS_WORLD_BODY) { bodies[j.body_a].kept = 1.0f; }
        if (j.body_b != TFS_WORLD_BODY) { bodies[j.body_b].kept = 1.0f; }
    }
    for (uint32_t k = 0; k < n; ++k) {
        const tfs_joint& j = joints[active[k]];
        tfj::End a = loadEnd(j.body_a, bodies, motion, world), b = loadEnd(j.body_b, bodies, motion, world);
        tfj::warmStartJoint(a, b, j);
        storeEndVelocity(j.body_a, a, motion, world);
        storeEndVelocity(j.body_b, b, motion, world);
    }
}

// solveJoints (src/joint.cpp:584-590): biased or relaxed, the relax pass adding each joint's reaction.
__global__ void k_joints_solve(tfs_joint* joints, const uint32_t* active, const DevCounts* cnt, const tfs_body* bodies,
                               float4* motion, float4* world, tfj::JCtx jc, int useBias) {
    if (gid() != 0) { return; }
    const uint32_t n = cnt->nJointsActive;
    for (uint32_t k = 0; k < n; ++k) {
        tfs_joint& j = joints[active[k]];
        tfj::End a = loadEnd(j.body_a, bodies, motion, world), b = loadEnd(j.body_b, bodies, motion, world);
        tfj::solveJoint(a, b, j, jc, useBias != 0);
        if (!useBias) { tfj::addReaction(j); }
        storeEndVelocity(j.body_a, a, motion, world);
        storeEndVelocity(j.body_b, b, motion, world);
    }
}

// After the final speed limits: breakJoints (src/world.cpp:1545-1563), then settleJoints over the joints
// left (src/joint.cpp:544-553, kSettleIterations = 2 rounds, src/world.cpp:16).
__global__ void k_joints_end(tfs_joint* joints, uint32_t* active, DevCounts* cnt, tfs_body* bodies,
                             const float4* motion, const float4* world, float dt) {
    if (gid() != 0) { return; }
    const uint32_t n = cnt->nJointsActive;
    uint32_t kept = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t ji = active[k];
        tfs_joint& j = joints[ji];
        const V3 force = tfj::vdiv(ld3(j.force), dt), torque = tfj::vdiv(ld3(j.torque), dt);
        st3(j.force, force);
        st3(j.torque, torque);
        if (length(force) > j.break_force || length(torque) > j.break_torque) {
            j.broken = 1;
            st3(j.linear_impulse, vec(0.0f, 0.0f, 0.0f));
            st3(j.angular_impulse, vec(0.0f, 0.0f, 0.0f));
            st3(j.friction_turn, vec(0.0f, 0.0f, 0.0f));
            j.axial_impulse = j.motor_impulse = j.lower_impulse = j.upper_impulse = 0.0f;
            j.friction_impulse = j.swing_impulse = 0.0f;
            continue;
        }
        active[kept++] = ji;
    }
    cnt->nJointsActive = kept;
    for (int round = 0; round < 2; ++round) {
        for (uint32_t k = 0; k < kept; ++k) {
            const tfs_joint& j = joints[active[k]];
            tfj::End a = loadEnd(j.body_a, bodies, motion, world), b = loadEnd(j.body_b, bodies, motion, world);
            tfj::settleJoint(a, b, j);
            storeEndPose(j.body_a, a, bodies);
            storeEndPose(j.body_b, b, bodies);
        }
    }
}

// Then, when the step had active joints, every awake dynamic body's inverse mass (src/world.cpp:1688-1690).
__global__ void k_joints_mass(tfs_body* bodies, uint32_t nb, const DevCounts* cnt) {
    if (cnt->nJointsBegun == 0) { return; }
    for (uint32_t i = gid(); i < nb; i += gstride()) {
        tfs_body& b = bodies[i];
        if (b.alive && b.type == TFS_DYNAMIC && b.awake) { updateInverseMass(b); }
    }
}

// Island ids are the lowest body index of the island; with several worlds, local to the body's world.
// Given `unsettled`, both run only if that round linked nothing (*unsettled == 0).
__global__ void k_island_time(tfs_body* bodies, uint32_t n, const uint32_t* parent, uint32_t* islandTime,
                              const uint32_t* bodyWorld, const uint32_t* worldStart, const uint32_t* unsettled) {
    if (unsettled && *unsettled != 0u) { return; }
    // Sleep times are never negative, so their bits order as the floats do. A large island fills whole
    // waves: its lanes take their minimum across the wave first and make one atomic between them.
    for (uint32_t base = blockIdx.x * blockDim.x; base < n; base += gstride()) {
        const uint32_t i = base + threadIdx.x;
        bool valid = false;
        uint32_t r = 0, t = 0xffffffffu;
        if (i < n) {
            tfs_body& b = bodies[i];
            if (b.alive && awakeDynamic(b)) {
                valid = true;
                r = parent[i];
                b.island = bodyWorld ? r - worldStart[bodyWorld[i]] : r;
                memcpy(&t, &b.sleep_time, 4);
            }
        }
        const unsigned long long act = __ballot(valid);
        if (act == 0ull) { continue; }
        const int first = __ffsll((long long)act) - 1;
        const uint32_t r0 = (uint32_t)__shfl((int)r, first);
        uint32_t m = (valid && r == r0) ? t : 0xffffffffu;
        for (int off = 32; off > 0; off >>= 1) { m = min(m, (uint32_t)__shfl_xor((int)m, off)); }
        if (valid && r != r0) { atomicMin(&islandTime[r], t); }
        if ((int)(threadIdx.x % 64) == first) { atomicMin(&islandTime[r0], m); }
    }
}

__global__ void k_sleep_apply(tfs_body* bodies, uint32_t n, const uint32_t* parent, const uint32_t* islandTime,
                              float timeToSleep, float4* motion, float4* mass, const uint32_t* unsettled) {
    if (unsettled && *unsettled != 0u) { return; }
    for (uint32_t i = gid(); i < n; i += gstride()) {
        tfs_body& b = bodies[i];
        if (!b.alive || !awakeDynamic(b)) { continue; }
        float t;
        memcpy(&t, &islandTime[parent[i]], 4);
        if (t < timeToSleep) { continue; }
        b.awake = 0;
        b.v[0] = b.v[1] = b.v[2] = 0.0f;
        b.w[0] = b.w[1] = b.w[2] = 0.0f;
        updateInverseMass(b);
        storeMotion(motion, i, Motion{ vec(0, 0, 0), vec(0, 0, 0), ld3(b.dp), ld4(b.dq) });
        storeMass(mass, i, b);
    }
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    for (uint32_t i = gid(); i < n; i += gstride()) { p[i] = v; }
}
__global__ void k_fill_u8(uint8_t* p, uint32_t n, uint8_t v) {
    for (uint32_t i = gid(); i < n; i += gstride()) { p[i] = v; }
}
__global__ void k_fill_u64(unsigned long long* p, uint32_t n) {
    for (uint32_t i = gid(); i < n; i += gstride()) { p[i] = 0ull; }
}

// ---- host side -----------------------------------------------------------------------------------

uint32_t blocksFor(uint32_t n, uint32_t perBlock = kThreads, uint32_t cap = 8192) {
    const uint32_t b = (n + perBlock - 1) / perBlock;
    return b < 1 ? 1 : (b > cap ? cap : b);
}

template <class T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
};

}  // namespace

struct tfs_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};   // step start, solve end, sleep end, refresh gather, substep loop start / end
    std::string error;
    tfs_settings settings{};
    uint32_t nb = 0, nc = 0, capB = 0, capC = 0;
    // independent worlds (tfs_set_worlds): body -> world, per-world refresh flags, host copy of the map
    uint32_t nWorlds = 1;
    uint32_t* bodyWorld = nullptr;
    uint8_t* refreshing = nullptr;
    uint32_t* worldStart = nullptr;
    uint32_t worldCap = 0;
    std::vector<uint32_t> hostBodyWorld;
    bool haveBodies = false, haveContacts = false;
    float solveMs = 0.0f, sleepMs = 0.0f;
    // hook-and-jump rounds of the island pass launched before the host first checks (1..kIslandAhead;
    // TWOFOLD_ISLAND_AHEAD lowers it, so that tests reach the host-driven rounds)
    int islandAhead = kIslandAhead;
    // tfs_step runs the sleep pass after the solve block (off: the caller runs it, tfs_sleep)
    bool stepSleep = true;
    // Measurement: HIP events around each row launch, summed per family (0 warm start, 1 solve,
    // 2 restitution), with the algorithmic bytes of the rows each launch covers.
    bool profiling = false;
    std::vector<hipEvent_t> evPool;
    struct Pending {
        int family;
        size_t ev;
        double bytes;
    };
    std::vector<Pending> pending;
    // families: 0 warm start, 1 solve, 2 restitution (event-timed launches); 3 waiting for the
    // refresh's contacts to reach the host, 4 the refresh callback (host clock); 5 the solve passes
    // over the rest lists (event-timed, all launches of one pass over one list together)
    uint64_t profLaunches[kProfileFamilies] = {};
    double profMs[kProfileFamilies] = {}, profBytes[kProfileFamilies] = {};
    uint32_t lastCounts[8] = {};

    // bodies
    tfs_body* bodies = nullptr;
    float4* motion = nullptr;
    float4* mass = nullptr;
    uint8_t* moving = nullptr;
    uint8_t* struck = nullptr;
    uint32_t* deg = nullptr;
    uint32_t* epStart = nullptr;
    uint32_t* cursor = nullptr;
    unsigned long long* word = nullptr;
    uint32_t* parent = nullptr;
    uint32_t* islandTime = nullptr;
    // contacts
    tfs_contact* contacts = nullptr;
    tfs_contact* fresh = nullptr;   // the refresh callback's manifolds, copied in
    uint32_t* fActive = nullptr;
    uint32_t* fFast = nullptr;
    uint32_t* fFastSolid = nullptr;
    uint32_t* active = nullptr;
    uint32_t* fast = nullptr;
    uint32_t* fastSolid = nullptr;
    uint32_t* fastOrder = nullptr;        // the fast lists' positions by collision cost (k_refresh_order)
    uint32_t* fastSolidOrder = nullptr;
    uint32_t* impactFlag = nullptr;
    uint32_t* impactPos = nullptr;
    uint32_t* impact = nullptr;
    uint32_t* rowsList = nullptr;
    uint32_t* impactRowsList = nullptr;
    uint2* edges = nullptr;               // sleep islands: one edge per contact, then one per joint
    // joints (tfs_upload_joints): the records, the active list of the step, the world body's v and w
    tfs_joint* joints = nullptr;
    uint32_t* jointActive = nullptr;
    float4* worldMotion = nullptr;
    uint32_t nJoints = 0, capJ = 0;
    uint32_t jointBodies = 0;             // 1 + the highest body index a resident joint names
    // the rest lists in levels (active rest, impact rest): scratch, and per list its order and level starts
    uint8_t* restTracked = nullptr;
    uint32_t *restDeg = nullptr, *restStart = nullptr, *restCursor = nullptr, *restEp = nullptr, *restRank = nullptr,
             *restLevel = nullptr, *restCount = nullptr;
    unsigned long long* restWord = nullptr;
    uint4* restEntries[2] = {};
    uint32_t* restLevelStart[2] = {};
    uint32_t restLevels[2] = {};
    uint32_t restWide[2] = {};     // leading levels wider than a rest work-group
    double restBytes[2] = {};      // algorithmic bytes of one solve pass over each rest list
    uint32_t *wrestDeg = nullptr, *wrestStart = nullptr, *wrestCursor = nullptr, *wrestEp = nullptr;  // rest warm start lists
    float4* restH = nullptr;        // the active rest contacts packed in level order (RestPack)
    float4* restP = nullptr;
    float* restPeak = nullptr;
    float4* restAx = nullptr;
    float4* restPre = nullptr;      // per pass: the planar rest contacts' normal-row biases and factors
    uint32_t* restPos = nullptr;
    uint8_t* colors = nullptr;
    uint32_t* ep = nullptr;
    uint32_t* rank = nullptr;
    uint32_t* rowOf = nullptr;
    // reportTouches: per contact, the pair's event and hit flags (set on its first contact) and its hit;
    // the compacted event list and hits
    uint32_t* touchFlag = nullptr;
    uint32_t* hitFlag = nullptr;
    uint32_t* touchList = nullptr;
    uint32_t* hitList = nullptr;
    tfs_hit* hitData = nullptr;
    tfs_hit* hits = nullptr;
    uint32_t touchCounts[3] = {};   // events, hits, touching contacts of the last step
    // the contact stream (tfs_stream_contacts): each contact's reuse data, and the stream's inputs, its
    // per-pair scratch and its output before it replaces the resident contacts
    tfs_contact_ext* ext = nullptr;
    tfs_contact* contactsNext = nullptr;
    tfs_contact_ext* extNext = nullptr;
    uint32_t streamCap = 0;
    tfs_pair* pairs = nullptr;
    uint32_t *pairLo = nullptr, *pairHi = nullptr, *pairKeep = nullptr, *pairStart = nullptr;
    uint32_t pairCap = 0;
    tfs_manifold* manifolds = nullptr;
    uint32_t manifoldCap = 0;
    uint32_t* hist = nullptr;
    size_t histCap = 0;
    unsigned long long* status = nullptr;  // look-back scan status words
    size_t statusCap = 0;
    uint32_t scanEpoch = 0;
    int scanSlot = 0;
    Rows rows{};
    uint4* impactHdr = nullptr;     // impact rows' headers, .w = their active row
    BarrierLines* bar = nullptr;   // grid barrier counters
    BarrierBase barrierBase{};
    uint32_t maxPassBlocks = 0;
    uint32_t passBlockLimit = 0;    // blocks of a grid-barrier pass: half of what the device holds at once
    bool usePass = true;
    bool coopLaunch = false;        // grid-barrier kernels through hipLaunchCooperativeKernel (launchGrid)
    bool passTail = true;
    uint32_t tailSegRows = kTailSegRows;   // widest colour the tail takes
    bool madeAxes = false;
    // A substep without a refresh and without rest lists in one launch (k_substep_fused), and the
    // substep loop of a step captured in a HIP graph and launched as one (both part of the pass shape)
    bool fused = false;
    bool useGraph = false;
    uint32_t fusedBlockLimit = 0;   // blocks of the fused launch: half of what the device holds at once
    hipGraphExec_t graphExec = nullptr;
    // Autotuning of the pass shape (grid blocks, tail colour width, axes loaded or made): while on,
    // tuning step s runs candidate (s + s / C) % C of the C shapes, and after `tuneRounds` rounds the
    // shape whose steps spent the least time in their solve passes (median) is kept. A step without
    // a pass does not count. Every shape gives the same results.
    struct PassShape {
        uint32_t blocks, tailRows;
        bool made;
        bool fused;
    };
    std::vector<PassShape> tuneShapes;
    std::vector<std::vector<float>> tuneMs;
    uint32_t tuneStep = 0, tuneRounds = 0;
    float tunePassMs = 0.0f;        // this step's solve passes (coloured and rest), event-timed
    bool tuning = false;
    // per-body lists of coloured rows for the warm start
    uint32_t* bdeg = nullptr;
    uint32_t* bstart = nullptr;
    uint32_t* bcursor = nullptr;
    uint32_t* bep = nullptr;
    DevCounts* cnt = nullptr;
    DevCounts* hostCnt = nullptr;
    // Device memory carved from a fixed array instead of allocated (the context of tfs_step_noalloc):
    // every buffer comes from [arena, arena + arenaCap) once, at set-up, and nothing is ever freed.
    char* arena = nullptr;
    size_t arenaCap = 0, arenaUsed = 0;
    // pinned host arrays of the refresh callback, compact in the order of the list: the list, its
    // contacts, their bodies (two per contact), and the fresh manifolds
    tfs_body* pinBodies = nullptr;
    tfs_contact* pinContacts = nullptr;
    uint32_t* pinAgain = nullptr;
    tfs_contact* pinFresh = nullptr;
    tfs_shape* shapes = nullptr;    // one per body: the device refresh (tfs_set_body_shapes), or none
    uint32_t nShapes = 0, shapesCap = 0;
    uint32_t pinCap = 0;
    // the fixed-contact model (tfs_fixed_*) and its adjoint (tfs_adjoint_*): the counts of the step
    // that fixed the contacts, one checkpoint per step taken, one step's tape, the adjoint state
    bool fixedOn = false;
    float fixedDt = 0.0f;
    uint32_t fixedSteps = 0, fixedMax = 0, fixedLeft = 0;
    DevCounts fixedCounts{};
    float4* ckBody = nullptr;
    float4* ckRow = nullptr;
    float4* tapeBuf = nullptr;
    Adjoint adj{};
    float* dlcom = nullptr;
    float* muContact = nullptr;
};

namespace {

int fail(tfs_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) { c->error = buf; }
    return code;
}

#define HIP_TRY(call)                                                                          \
    do {                                                                                       \
        const hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) { return fail(ctx, TFS_E_DEVICE, "%s: %s", #call, hipGetErrorString(e_)); } \
    } while (0)

#define LAUNCH(kernel, grid, block, ...)                                                        \
    do {                                                                                       \
        hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, ctx->stream, __VA_ARGS__);      \
        const hipError_t e_ = hipGetLastError();                                               \
        if (e_ != hipSuccess) { return fail(ctx, TFS_E_DEVICE, "%s: %s", #kernel, hipGetErrorString(e_)); } \
    } while (0)

// Takes `bytes` from the context's fixed array (256-byte aligned), for a context that never allocates.
int carve(tfs_ctx* ctx, void** p, size_t bytes) {
    const size_t at = (ctx->arenaUsed + 255) & ~(size_t)255;
    if (at + bytes > ctx->arenaCap) {
        return fail(ctx, TFS_E_CAPACITY, "the fixed scratch of %zu bytes is too small (%zu more needed)", ctx->arenaCap,
                    at + bytes - ctx->arenaCap);
    }
    *p = ctx->arena + at;
    ctx->arenaUsed = at + bytes;
    return 0;
}

template <class T>
int grow(tfs_ctx* ctx, T*& p, size_t n) {
    if (ctx->arena) { return carve(ctx, (void**)&p, sizeof(T) * (n ? n : 1)); }
    if (p) { HIP_TRY(hipFree(p)); }
    p = nullptr;
    HIP_TRY(hipMalloc((void**)&p, sizeof(T) * (n ? n : 1)));
    return 0;
}

int ensureScratch(tfs_ctx* ctx);

int allocBodies(tfs_ctx* ctx, uint32_t cap) {
    int r;
    if ((r = grow(ctx, ctx->bodies, cap)) || (r = grow(ctx, ctx->motion, 4 * (size_t)cap)) ||
        (r = grow(ctx, ctx->mass, 6 * (size_t)cap)) || (r = grow(ctx, ctx->moving, cap)) ||
        (r = grow(ctx, ctx->struck, cap)) || (r = grow(ctx, ctx->deg, cap)) || (r = grow(ctx, ctx->epStart, cap)) ||
        (r = grow(ctx, ctx->cursor, cap)) || (r = grow(ctx, ctx->word, cap)) || (r = grow(ctx, ctx->parent, cap)) ||
        (r = grow(ctx, ctx->islandTime, cap)) || (r = grow(ctx, ctx->bdeg, cap)) || (r = grow(ctx, ctx->bstart, cap)) ||
        (r = grow(ctx, ctx->bcursor, cap)) || (r = grow(ctx, ctx->bodyWorld, cap)) ||
        (r = grow(ctx, ctx->restTracked, cap)) || (r = grow(ctx, ctx->restDeg, cap)) ||
        (r = grow(ctx, ctx->restStart, cap)) || (r = grow(ctx, ctx->restCursor, cap)) ||
        (r = grow(ctx, ctx->restWord, cap)) || (r = grow(ctx, ctx->wrestDeg, cap)) ||
        (r = grow(ctx, ctx->wrestStart, cap)) || (r = grow(ctx, ctx->wrestCursor, cap))) {
        return r;
    }
    ctx->capB = cap;
    return ensureScratch(ctx);
}

int allocRows(tfs_ctx* ctx, Rows& r, uint32_t cap) {
    int e;
    if ((e = grow(ctx, r.hdr, cap)) || (e = grow(ctx, r.nrm, cap)) || (e = grow(ctx, r.tg1, cap)) ||
        (e = grow(ctx, r.tg2, cap)) || (e = grow(ctx, r.pa, 4 * (size_t)cap)) || (e = grow(ctx, r.pb, 4 * (size_t)cap)) ||
        (e = grow(ctx, r.imp, 4 * (size_t)cap)) || (e = grow(ctx, r.peak, 4 * (size_t)cap)) ||
        (e = grow(ctx, r.ax, 60 * (size_t)cap)) || (e = grow(ctx, r.sep, cap))) {
        return e;
    }
    r.cap = cap;
    return 0;
}

int allocContacts(tfs_ctx* ctx, uint32_t cap) {
    int r;
    if (cap >= (1u << (32 - kRestPosShift))) {
        return fail(ctx, TFS_E_ARG, "contact capacity %u is beyond the limit of %u", cap, (1u << (32 - kRestPosShift)) - 1u);
    }
    if ((r = grow(ctx, ctx->contacts, cap)) || (r = grow(ctx, ctx->fresh, cap)) || (r = grow(ctx, ctx->fActive, cap)) ||
        (r = grow(ctx, ctx->fFast, cap)) || (r = grow(ctx, ctx->fFastSolid, cap)) || (r = grow(ctx, ctx->active, cap)) ||
        (r = grow(ctx, ctx->fast, cap)) || (r = grow(ctx, ctx->fastSolid, cap)) ||
        (r = grow(ctx, ctx->fastOrder, std::min(cap, kOrderMax))) ||
        (r = grow(ctx, ctx->fastSolidOrder, std::min(cap, kOrderMax))) ||
        (r = grow(ctx, ctx->impactFlag, cap)) || (r = grow(ctx, ctx->impactPos, cap)) ||
        (r = grow(ctx, ctx->impact, cap)) || (r = grow(ctx, ctx->rowsList, cap)) ||
        (r = grow(ctx, ctx->impactRowsList, cap)) || (r = grow(ctx, ctx->colors, cap)) ||
        (r = grow(ctx, ctx->ep, 2 * (size_t)cap)) || (r = grow(ctx, ctx->rank, 2 * (size_t)cap)) ||
        (r = grow(ctx, ctx->rowOf, cap)) || (r = grow(ctx, ctx->edges, (size_t)cap + ctx->capJ)) ||
        (r = grow(ctx, ctx->restEp, 2 * (size_t)cap)) || (r = grow(ctx, ctx->restRank, 2 * (size_t)cap)) ||
        (r = grow(ctx, ctx->restLevel, cap)) || (r = grow(ctx, ctx->restCount, cap)) ||
        (r = grow(ctx, ctx->restEntries[0], cap)) || (r = grow(ctx, ctx->restEntries[1], cap)) ||
        (r = grow(ctx, ctx->restLevelStart[0], (size_t)cap + 1)) ||
        (r = grow(ctx, ctx->restLevelStart[1], (size_t)cap + 1)) || (r = grow(ctx, ctx->restH, 5 * (size_t)cap)) ||
        (r = grow(ctx, ctx->restP, 12 * (size_t)cap)) || (r = grow(ctx, ctx->restPeak, 4 * (size_t)cap)) || (r = grow(ctx, ctx->restAx, 8 * (size_t)cap)) ||
        (r = grow(ctx, ctx->restPre, 3 * (size_t)cap)) ||
        (r = grow(ctx, ctx->restPos, cap)) || (r = grow(ctx, ctx->wrestEp, 2 * (size_t)cap)) ||
        (r = grow(ctx, ctx->touchFlag, cap)) || (r = grow(ctx, ctx->hitFlag, cap)) || (r = grow(ctx, ctx->touchList, cap)) ||
        (r = grow(ctx, ctx->hitList, cap)) || (r = grow(ctx, ctx->hitData, cap)) || (r = grow(ctx, ctx->hits, cap)) ||
        (r = grow(ctx, ctx->ext, cap))) {
        return r;
    }
    if ((r = allocRows(ctx, ctx->rows, cap)) || (r = grow(ctx, ctx->impactHdr, cap)) || (r = grow(ctx, ctx->bep, 2 * (size_t)cap))) {
        return r;
    }
    ctx->capC = cap;
    return ensureScratch(ctx);
}

// The row-order histogram and the scans' status words, sized for the body and contact capacities so
// that a step never allocates.
int ensureScratch(tfs_ctx* ctx) {
    int r;
    const size_t hist = (size_t)kKeys * blocksFor(ctx->capC, kThreads, 1u << 30);
    if (hist > ctx->histCap) {
        if ((r = grow(ctx, ctx->hist, hist))) { return r; }
        ctx->histCap = hist;
    }
    const size_t most = std::max<size_t>(std::max<size_t>(ctx->capB, ctx->capC), hist);
    const size_t words = 3 * std::max<size_t>(1, (most + kScanTile - 1) / kScanTile);
    if (words > ctx->statusCap) {
        if ((r = grow(ctx, ctx->status, words))) { return r; }
        LAUNCH(k_fill_u64, blocksFor((uint32_t)words), kThreads, ctx->status, (uint32_t)words);
        ctx->statusCap = words;
        ctx->scanEpoch = 0;
    }
    return 0;
}

// Scans of up to three arrays of at most nMax items each, in one launch. Each launch of a step takes
// its own ticket counter (cleared with the step's counters) and a fresh epoch for its status words.
int scanJobs(tfs_ctx* ctx, const ScanJob* job, int count, uint32_t nMax) {
    const uint32_t tiles = std::max<uint32_t>(1, (nMax + kScanTile - 1) / kScanTile);
    if ((size_t)tiles * 3 > ctx->statusCap) { return fail(ctx, TFS_E_STATE, "scan of %u items beyond capacity", nMax); }
    if (ctx->scanSlot >= kScanSlots) { return fail(ctx, TFS_E_DEVICE, "more than %d scans in one step", kScanSlots); }
    if (++ctx->scanEpoch >= kEpochLimit) {
        LAUNCH(k_fill_u64, blocksFor((uint32_t)ctx->statusCap), kThreads, ctx->status, (uint32_t)ctx->statusCap);
        ctx->scanEpoch = 1;
    }
    ScanJobs jobs{};
    for (int k = 0; k < count; ++k) { jobs.job[k] = job[k]; }
    jobs.tiles = tiles;
    LAUNCH(k_scan, tiles * (uint32_t)count, kScanThreads, jobs, ctx->status, ctx->scanEpoch,
           &ctx->cnt->scanTicket[ctx->scanSlot++], ctx->cnt);
    return 0;
}

// Exclusive scan of in[0..n) (n read from *nPtr on the device when given, else nMax); optionally
// the prefix array and/or the compacted index list; the total goes to *total.
int scan(tfs_ctx* ctx, const uint32_t* in, const uint32_t* nPtr, uint32_t nMax, uint32_t* out, uint32_t* list,
         uint32_t* total) {
    const ScanJob job{ in, nPtr, nMax, out, list, total };
    return scanJobs(ctx, &job, 1, nMax);
}

// Stable row order of `list` (length *nPtr <= nMax) into out; colour starts into starts[0..32].
int rowOrder(tfs_ctx* ctx, const uint32_t* list, const uint32_t* nPtr, uint32_t nMax, uint32_t* out, uint32_t* starts,
             uint32_t* points) {
    const uint32_t nBlocks = blocksFor(nMax, kThreads, 1u << 30);
    if ((size_t)kKeys * nBlocks > ctx->histCap) { return fail(ctx, TFS_E_STATE, "row order beyond capacity"); }
    int r;
    LAUNCH(k_row_hist, nBlocks, kThreads, list, nPtr, ctx->contacts, ctx->colors, ctx->hist, nBlocks);
    if ((r = scan(ctx, ctx->hist, nullptr, kKeys * nBlocks, ctx->hist, nullptr, &ctx->cnt->scanTotal))) { return r; }
    LAUNCH(k_row_scatter, nBlocks, kThreads, list, nPtr, ctx->contacts, ctx->colors, ctx->hist, nBlocks, out);
    LAUNCH(k_row_starts, 1, 64, ctx->hist, nBlocks, starts, points);
    return 0;
}

// Algorithmic bytes of one pass over a contact with n points, from the reference's row model
// (src/internal.h:579-605): contact data, two bodies gathered and scattered, and per point the three
// axes, anchors, separation and impulses.
double rowBytes(int mode, uint32_t rows, uint32_t points) {
    switch (mode) {
        case kWarm: return 224.0 * rows + 244.0 * points;
        case kRestitute: return 224.0 * rows + 96.0 * points;
        default: return 224.0 * rows + 296.0 * points;
    }
}

// Algorithmic bytes of one solve pass over rest contacts, from the same row model: a planar contact
// with n points 248 + 68n (2D axes, planar inverse mass and inertia), a 3D one 224 + 296n.
double restPassBytes(uint32_t planar, uint32_t planarPoints, uint32_t solid, uint32_t solidPoints) {
    return 248.0 * planar + 68.0 * planarPoints + 224.0 * solid + 296.0 * solidPoints;
}

int runRows(tfs_ctx* ctx, int mode, const uint4* at, const uint32_t* starts, const uint32_t* points, const Ctx& c,
            const Tape& tape = Tape{}) {
    for (int col = 0; col < TFS_COLORS; ++col) {
        const uint32_t b = starts[col], e = starts[col + 1];
        if (e <= b) { continue; }
        const uint32_t grid = (e - b + kThreads - 1) / kThreads;
        size_t ev = 0;
        if (ctx->profiling) {
            ev = ctx->pending.size() * 2;
            while (ctx->evPool.size() < ev + 2) {
                hipEvent_t x;
                HIP_TRY(hipEventCreate(&x));
                ctx->evPool.push_back(x);
            }
            const int family = mode == kWarm ? 0 : mode == kRestitute ? 2 : 1;
            ctx->pending.push_back({ family, ev, rowBytes(mode, e - b, points[col]) });
            HIP_TRY(hipEventRecord(ctx->evPool[ev], ctx->stream));
        }
        const float4* mass = ctx->mass;
        if (ctx->madeAxes) {
            switch (mode) {
                case kBias: LAUNCH((k_rows<kBias, true>), grid, kThreads, ctx->rows, at, b, e, ctx->motion, mass, c, tape); break;
                case kRelax: LAUNCH((k_rows<kRelax, true>), grid, kThreads, ctx->rows, at, b, e, ctx->motion, mass, c, tape); break;
                default: LAUNCH((k_rows<kRestitute, true>), grid, kThreads, ctx->rows, at, b, e, ctx->motion, mass, c, tape); break;
            }
        } else {
            switch (mode) {
                case kBias: LAUNCH((k_rows<kBias, false>), grid, kThreads, ctx->rows, at, b, e, ctx->motion, mass, c, tape); break;
                case kRelax: LAUNCH((k_rows<kRelax, false>), grid, kThreads, ctx->rows, at, b, e, ctx->motion, mass, c, tape); break;
                default: LAUNCH((k_rows<kRestitute, false>), grid, kThreads, ctx->rows, at, b, e, ctx->motion, mass, c, tape); break;
            }
        }
        if (ctx->profiling) { HIP_TRY(hipEventRecord(ctx->evPool[ev + 1], ctx->stream)); }
    }
    return 0;
}

// A kernel with a grid barrier (k_pass, k_rest_levels, k_pass_adj): every block must be resident at
// once. Its grid holds at most passBlockLimit blocks (half of what the device holds at once), and a
// stream starts a kernel only when the one before it has finished, so when nothing else runs on the
// device the blocks are resident together whatever the launch. The plain launch (the default) spares
// the queue barriers the cooperative launch puts around the kernel (~6 us on each side: 5-7% of a
// step); the cooperative one (tfs_set_cooperative, or TWOFOLD_COOP=on) also holds when other work
// shares the device. A barrier whose blocks never all arrive ends at its spin limit with an error.
int launchGrid(tfs_ctx* ctx, const void* fn, uint32_t grid, void** args) {
    if (ctx->coopLaunch) {
        HIP_TRY(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(kPassThreads), args, 0, ctx->stream));
    } else {
        HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(kPassThreads), args, 0, ctx->stream));
    }
    return 0;
}

// A pass over the listed row sets (active colours, then impact colours when `withImpact`) in one
// launch with a grid barrier between colours (launchGrid).
// The plan of one pass over the listed row sets (active colours, then impact colours when
// `withImpact`), its grid and its algorithmic bytes; plan.n == 0 when there are no rows.
void planPass(const tfs_ctx* ctx, int mode, const DevCounts& hc, bool withActive, bool withImpact, PassPlan& plan,
              uint32_t& grid, double& bytes) {
    plan = PassPlan{};
    uint32_t widest = 0;
    bytes = 0.0;
    auto add = [&](const uint32_t* starts, const uint32_t* points) {
        for (int col = 0; col < TFS_COLORS; ++col) {
            if (starts[col + 1] <= starts[col]) { continue; }
            plan.begin[plan.n] = starts[col];
            plan.end[plan.n] = starts[col + 1];
            widest = std::max(widest, starts[col + 1] - starts[col]);
            bytes += rowBytes(mode, starts[col + 1] - starts[col], points[col]);
            ++plan.n;
        }
    };
    if (withActive) { add(hc.colorStart, hc.colorPoints); }
    plan.impactFrom = plan.n;
    // the separations of every coloured active row, which the impact rows of the pass read too (an
    // impact-only launch follows the active launch of the same pass, which computed them)
    plan.sepRows = withActive && plan.n > 0 ? hc.colorStart[TFS_COLORS] : 0u;
    if (withImpact) { add(hc.impactColorStart, hc.impactColorPoints); }
    grid = 1;
    if (plan.n == 0) { return; }
    // The tail: the trailing colours of at most one wave of rows each, at most one block of rows in
    // all (the active rows and the impact visits), when there are at least two of them.
    uint32_t tailFrom = plan.n, tailRows = 0;
    while (ctx->passTail && tailFrom > 0) {
        const uint32_t w = plan.end[tailFrom - 1] - plan.begin[tailFrom - 1];
        if (w > ctx->tailSegRows || tailRows + w > (uint32_t)kPassThreads) { break; }
        tailRows += w;
        --tailFrom;
    }
    if (plan.n - tailFrom < 2) { tailFrom = plan.n; }
    plan.tailFrom = tailFrom;
    widest = 0;
    for (uint32_t s = 0; s < tailFrom; ++s) { widest = std::max(widest, plan.end[s] - plan.begin[s]); }
    grid = std::max<uint32_t>(1, std::min<uint32_t>((widest + 63) / 64, ctx->maxPassBlocks));
}

int runPass(tfs_ctx* ctx, int mode, const DevCounts& hc, bool withActive, bool withImpact, const Ctx& c,
            const Tape& tape = Tape{}) {
    PassPlan plan;
    uint32_t grid = 1;
    double bytes = 0.0;
    planPass(ctx, mode, hc, withActive, withImpact, plan, grid, bytes);
    if (plan.n == 0) { return 0; }
    const uint32_t barriers = passBarriers(plan);
    size_t ev = 0;
    const bool timed = ctx->profiling;
    if (timed) {
        ev = ctx->pending.size() * 2;
        while (ctx->evPool.size() < ev + 2) {
            hipEvent_t x;
            HIP_TRY(hipEventCreate(&x));
            ctx->evPool.push_back(x);
        }
        ctx->pending.push_back({ 1, ev, bytes });
        HIP_TRY(hipEventRecord(ctx->evPool[ev], ctx->stream));
    }
    const uint4* at = ctx->impactHdr;
    const BarrierBase base = ctx->barrierBase;
    const float4* mass = ctx->mass;
    void* args[] = { &ctx->rows, (void*)&at, &plan, &ctx->motion, (void*)&mass, (void*)&c, &ctx->bar, (void*)&base, &ctx->cnt,
                     (void*)&tape };
    const bool rec = tape.v != nullptr;
    const void* fn = ctx->madeAxes
                         ? (mode == kBias ? (rec ? (const void*)k_pass<kBias, true, true> : (const void*)k_pass<kBias, true, false>)
                                          : (rec ? (const void*)k_pass<kRelax, true, true> : (const void*)k_pass<kRelax, true, false>))
                         : (mode == kBias ? (rec ? (const void*)k_pass<kBias, false, true> : (const void*)k_pass<kBias, false, false>)
                                          : (rec ? (const void*)k_pass<kRelax, false, true> : (const void*)k_pass<kRelax, false, false>));
    if (int e = launchGrid(ctx, fn, grid, args)) { return e; }
    advanceBarrier(ctx->barrierBase, grid, barriers);
    if (timed) { HIP_TRY(hipEventRecord(ctx->evPool[ev + 1], ctx->stream)); }
    return 0;
}

int runRest(tfs_ctx* ctx, int mode, const uint32_t* list, uint32_t begin, uint32_t end, const Ctx& c);

// One solve pass, biased or relax: the coloured active rows, the active rest in order, then the same
// for the impact rows (src/world.cpp:1675-1680; src/lanes.cpp:419-424).
int solvePass(tfs_ctx* ctx, int mode, const DevCounts& hc, uint32_t laned, uint32_t impactLaned, const Ctx& c) {
    int r;
    const bool restA = hc.nActive > laned, impact = hc.nImpact > 0;
    if (!ctx->usePass) {
        if ((r = runRows(ctx, mode, nullptr, hc.colorStart, hc.colorPoints, c)) ||
            (r = runRest(ctx, mode, ctx->rowsList, laned, hc.nActive, c))) {
            return r;
        }
        if (impact && ((r = runRows(ctx, mode, ctx->impactHdr, hc.impactColorStart, hc.impactColorPoints, c)) ||
                       (r = runRest(ctx, mode, ctx->impactRowsList, impactLaned, hc.nImpact, c)))) {
            return r;
        }
        return 0;
    }
    if (restA) {
        if ((r = runPass(ctx, mode, hc, true, false, c)) || (r = runRest(ctx, mode, ctx->rowsList, laned, hc.nActive, c))) {
            return r;
        }
        if (impact && (r = runPass(ctx, mode, hc, false, true, c))) { return r; }
    } else if ((r = runPass(ctx, mode, hc, true, impact, c))) {
        return r;
    }
    if (impact && (r = runRest(ctx, mode, ctx->impactRowsList, impactLaned, hc.nImpact, c))) { return r; }
    return 0;
}

// Packs the rows and builds the impact rows and each body's list of rows.
int packRows(tfs_ctx* ctx, uint32_t laned, uint32_t impactLaned, bool lists) {
    LAUNCH(k_pack, blocksFor(laned), kThreads, ctx->rowsList, laned, ctx->contacts, ctx->moving, ctx->mass, ctx->rows,
           ctx->rowOf);
    if (impactLaned > 0) {
        LAUNCH(k_impact_rows, blocksFor(impactLaned), kThreads, ctx->impactRowsList, impactLaned, ctx->rowOf,
               ctx->rows.hdr, ctx->impactHdr);
    }
    if (!lists) { return 0; }
    const uint32_t nb = ctx->nb;
    LAUNCH(k_bep_count, blocksFor(laned), kThreads, laned, ctx->rows.hdr, ctx->bdeg);
    int r;
    if ((r = scan(ctx, ctx->bdeg, nullptr, nb, ctx->bstart, nullptr, &ctx->cnt->scanTotal))) { return r; }
    LAUNCH(k_bep_fill, blocksFor(laned), kThreads, laned, ctx->rows.hdr, ctx->bstart, ctx->bcursor, ctx->bep);
    LAUNCH(k_bep_sort, blocksFor(nb), kThreads, nb, ctx->bdeg, ctx->bstart, ctx->bep);
    return 0;
}

int collectProfile(tfs_ctx* ctx) {
    for (const auto& p : ctx->pending) {
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, ctx->evPool[p.ev], ctx->evPool[p.ev + 1]));
        if (!ctx->profiling) { continue; }
        ctx->profLaunches[p.family] += 1;
        ctx->profMs[p.family] += ms;
        ctx->profBytes[p.family] += p.bytes;
    }
    ctx->pending.clear();
    return 0;
}

// The levels of a rest list (list[0..n), the active or the impact rest), built once per step.
int buildRestLevels(tfs_ctx* ctx, int which, const uint32_t* list, uint32_t n) {
    if (n == 0) { return 0; }
    const uint32_t nb = ctx->nb;
    const RestScratch z{ ctx->restTracked, ctx->restDeg, ctx->restStart, ctx->restCursor, ctx->restEp, ctx->restRank,
                         ctx->restLevel, ctx->restCount, ctx->restWord };
    int r;
    LAUNCH(k_rest_track, blocksFor(nb), kThreads, nb, ctx->moving, z);
    LAUNCH(k_rest_roll, blocksFor(n), kThreads, list, n, ctx->contacts, z);
    LAUNCH(k_rest_ep_count, blocksFor(n), kThreads, list, n, ctx->contacts, z);
    if ((r = scan(ctx, z.deg, nullptr, nb, z.start, nullptr, &ctx->cnt->scanTotal))) { return r; }
    LAUNCH(k_rest_ep_fill, blocksFor(n), kThreads, list, n, ctx->contacts, z);
    LAUNCH(k_ep_rank, blocksFor(nb), kThreads, nb, z.deg, z.start, z.ep, z.rank);
    LAUNCH(k_rest_level, blocksFor(n, kColorThreads, 2048), kColorThreads, list, n, ctx->contacts, z, &ctx->cnt->restTicket[which],
           &ctx->cnt->restLevels[which], ctx->cnt);
    // level starts: the exclusive prefix of the level sizes, and n after the last level
    if ((r = scan(ctx, z.count, nullptr, n, ctx->restLevelStart[which], nullptr, ctx->restLevelStart[which] + n))) {
        return r;
    }
    LAUNCH(k_rest_place, blocksFor(n), kThreads, list, n, ctx->contacts, ctx->moving, z, ctx->restLevelStart[which],
           ctx->restEntries[which], ctx->cnt->restShape[which]);
    LAUNCH(k_rest_narrow, blocksFor(n), kThreads, ctx->restLevelStart[which], &ctx->cnt->restLevels[which],
           (uint32_t)kRestBlockThreads, &ctx->cnt->restNarrow[which]);
    return 0;
}

// A rest pass: the list's leading wide levels (more contacts than a rest work-group has threads) in
// one launch over many blocks with the grid barrier, the levels from the first narrow one on in one work-group.
RestPack restPack(const tfs_ctx* ctx) {
    return RestPack{ ctx->restH, ctx->restP, ctx->restPeak, ctx->restAx, ctx->restPre, ctx->restPos, ctx->capC };
}

// The start of a substep: velocities integrated and the warm start of every row, coloured and rest,
// one launch over the bodies (k_substep_start). Profile family 0 times it, with the warm start's bytes.
int substepStart(tfs_ctx* ctx, const Ctx& c, uint32_t laned, uint32_t points, bool rest, int parts) {
    size_t ev = 0;
    const bool timed = ctx->profiling && (parts & 2);
    if (timed) {
        ev = ctx->pending.size() * 2;
        while (ctx->evPool.size() < ev + 2) {
            hipEvent_t x;
            HIP_TRY(hipEventCreate(&x));
            ctx->evPool.push_back(x);
        }
        ctx->pending.push_back({ 0, ev, rowBytes(kWarm, laned, points) });
        HIP_TRY(hipEventRecord(ctx->evPool[ev], ctx->stream));
    }
    const uint32_t nb = ctx->nb;
    if (rest) {
        LAUNCH(k_substep_start<true>, blocksFor(nb, kBodyThreads), kBodyThreads, ctx->bodies, nb, ctx->motion, c, ctx->bdeg, ctx->bstart,
               ctx->bep, ctx->rows, ctx->wrestDeg, ctx->wrestStart, ctx->wrestEp, ctx->restEntries[0], restPack(ctx),
               ctx->mass, parts);
    } else {
        LAUNCH(k_substep_start<false>, blocksFor(nb, kBodyThreads), kBodyThreads, ctx->bodies, nb, ctx->motion, c, ctx->bdeg, ctx->bstart,
               ctx->bep, ctx->rows, nullptr, nullptr, nullptr, nullptr, RestPack{}, ctx->mass, parts);
    }
    if (timed) { HIP_TRY(hipEventRecord(ctx->evPool[ev + 1], ctx->stream)); }
    return 0;
}

int runRest(tfs_ctx* ctx, int mode, const uint32_t* list, uint32_t begin, uint32_t end, const Ctx& c) {
    if (end <= begin) { return 0; }
    const int which = list == ctx->impactRowsList ? 1 : 0;
    const uint32_t levels = ctx->restLevels[which], n = end - begin;
    if (levels == 0) { return fail(ctx, TFS_E_STATE, "rest list without levels"); }
    const uint32_t wide = std::min(ctx->restWide[which], levels);
    const uint4* entries = ctx->restEntries[which];
    const uint32_t* starts = ctx->restLevelStart[which];
    BarrierLines* bar = ctx->bar;
    const BarrierBase base = ctx->barrierBase;
    const float4* mass = ctx->mass;
    const RestPack rp = restPack(ctx);
    const bool viaPos = which == 1;
    size_t ev = 0;
    // Profile family 5 times the passes over the active rest list (the impact rest list's are a few
    // contacts each and would only dilute the per-pass figure).
    const bool timed = ctx->profiling && which == 0 && (mode == kBias || mode == kRelax);
    if (timed) {
        ev = ctx->pending.size() * 2;
        while (ctx->evPool.size() < ev + 2) {
            hipEvent_t x;
            HIP_TRY(hipEventCreate(&x));
            ctx->evPool.push_back(x);
        }
        ctx->pending.push_back({ 5, ev, ctx->restBytes[which] });
        HIP_TRY(hipEventRecord(ctx->evPool[ev], ctx->stream));
    }
    if (which == 0 && mode == kBias) {
        LAUNCH(k_rest_pre<true>, blocksFor(n), kThreads, entries, n, rp, ctx->motion, c);
    } else if (which == 0 && mode == kRelax) {
        LAUNCH(k_rest_pre<false>, blocksFor(n), kThreads, entries, n, rp, ctx->motion, c);
    }
    if (wide > 0) {
        const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((n + 63) / 64, ctx->maxPassBlocks));
        void* args[] = { (void*)&entries, (void*)&starts, (void*)&wide, (void*)&rp, (void*)&viaPos, &ctx->motion,
                         (void*)&mass, (void*)&c, (void*)&bar, (void*)&base, &ctx->cnt };
        const void* fn = mode == kBias ? (const void*)k_rest_levels<kBias, false>
                         : mode == kRelax ? (const void*)k_rest_levels<kRelax, false>
                                          : (const void*)k_rest_levels<kRestitute, false>;
        if (int e = launchGrid(ctx, fn, grid, args)) { return e; }
        advanceBarrier(ctx->barrierBase, grid, wide - 1);
    }
    if (wide < levels && which == 0 && mode != kRestitute && levels - wide <= kWalkMaxLevels) {
        const uint32_t* from = starts + wide;
        const uint32_t left = levels - wide;
        if (mode == kBias) {
            LAUNCH(k_rest_walk<kBias>, 1, kWalkThreads, entries, from, left, rp, ctx->motion, mass, c);
        } else {
            LAUNCH(k_rest_walk<kRelax>, 1, kWalkThreads, entries, from, left, rp, ctx->motion, mass, c);
        }
    } else if (wide < levels) {
        const uint32_t* from = starts + wide;
        const uint32_t left = levels - wide;
        switch (mode) {
            case kBias: LAUNCH((k_rest_levels<kBias, true>), 1, kRestBlockThreads, entries, from, left, rp, viaPos, ctx->motion, mass, c, bar, base, ctx->cnt); break;
            case kRelax: LAUNCH((k_rest_levels<kRelax, true>), 1, kRestBlockThreads, entries, from, left, rp, viaPos, ctx->motion, mass, c, bar, base, ctx->cnt); break;
            default: LAUNCH((k_rest_levels<kRestitute, true>), 1, kRestBlockThreads, entries, from, left, rp, viaPos, ctx->motion, mass, c, bar, base, ctx->cnt); break;
        }
    }
    if (timed) { HIP_TRY(hipEventRecord(ctx->evPool[ev + 1], ctx->stream)); }
    return 0;
}

// The error bits kernels raise: 16 a pair the device narrowphase does not collide, the others a
// bounded wait that ran out.
int kernelFail(tfs_ctx* ctx, uint32_t bits) {
    if (bits & 16u) {
        return fail(ctx, TFS_E_KERNEL, "the device refresh met two boxes that both stay extruded along z (error %u)", bits);
    }
    if (bits & 32u) {
        return fail(ctx, TFS_E_ARG, "a contact names a body out of range or has more than 4 points, or a shape is not "
                                    "a sphere, box or half-space of finite positive size (error %u)", bits);
    }
    return fail(ctx, TFS_E_KERNEL, "a kernel's bounded wait ran out (error %u)", bits);
}

int syncCounts(tfs_ctx* ctx) {
    HIP_TRY(hipMemcpyAsync(ctx->hostCnt, ctx->cnt, sizeof(DevCounts), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->hostCnt->error) { return kernelFail(ctx, ctx->hostCnt->error); }
    return 0;
}

template <class T>
int growPinned(tfs_ctx* ctx, T*& p, size_t n) {
    if (p) { HIP_TRY(hipHostFree(p)); }
    p = nullptr;
    HIP_TRY(hipHostMalloc((void**)&p, sizeof(T) * (n ? n : 1), hipHostMallocDefault));
    return 0;
}

int refreshApply(tfs_ctx* ctx, const uint32_t* again, uint32_t nAgain, uint32_t laned, uint32_t impactLaned, uint32_t nRest);

int refresh(tfs_ctx* ctx, int substep, bool half, uint32_t nAgain, float remaining, tfs_refresh_fn cb, void* user,
            const DevCounts& hc) {
    const uint32_t nb = ctx->nb;
    const uint32_t laned = hc.colorStart[TFS_COLORS], impactLaned = hc.impactColorStart[TFS_COLORS];
    const uint32_t* again = half ? ctx->fast : ctx->fastSolid;
    const uint32_t* againFlag = half ? ctx->fFast : ctx->fFastSolid;
    const uint8_t* refreshing = nullptr;
    const uint32_t* bodyWorld = nullptr;
    if (ctx->nWorlds > 1) {
        LAUNCH(k_fill_u8, blocksFor(ctx->nWorlds), kThreads, ctx->refreshing, ctx->nWorlds, (uint8_t)0);
        LAUNCH(k_refresh_worlds, blocksFor(nAgain), kThreads, again, nAgain, ctx->contacts, ctx->bodyWorld,
               ctx->refreshing);
        refreshing = ctx->refreshing;
        bodyWorld = ctx->bodyWorld;
    }
    const uint32_t nRest = hc.nActive - laned;
    if (nRest > 0) {
        LAUNCH(k_rest_unpack, blocksFor(nRest), kThreads, ctx->restEntries[0], nRest, restPack(ctx), ctx->contacts);
    }
    LAUNCH(k_refresh_pre, blocksFor(hc.nActive), kThreads, ctx->active, ctx->cnt, againFlag, ctx->contacts, ctx->rows,
           ctx->rowOf, ctx->colors, ctx->motion, refreshing, bodyWorld);
    LAUNCH(k_refresh_bodies, blocksFor(nb, kBodyThreads), kBodyThreads, ctx->bodies, nb, ctx->motion, ctx->mass, refreshing, bodyWorld);
    if (ctx->nShapes) {
        // the pairs collided on the device: no transfer, no callback
        LAUNCH(k_pack, blocksFor(laned), kThreads, ctx->rowsList, laned, ctx->contacts, ctx->moving, ctx->mass, ctx->rows,
               ctx->rowOf);
        size_t ev = 0;
        if (ctx->profiling) {
            ev = ctx->pending.size() * 2;
            while (ctx->evPool.size() < ev + 2) {
                hipEvent_t x;
                HIP_TRY(hipEventCreate(&x));
                ctx->evPool.push_back(x);
            }
            ctx->pending.push_back({ 4, ev, (double)nAgain });
            HIP_TRY(hipEventRecord(ctx->evPool[ev], ctx->stream));
        }
        const uint32_t* order = nAgain > kOrderMax ? nullptr : half ? ctx->fastOrder : ctx->fastSolidOrder;
        LAUNCH(k_refresh_collide, blocksFor(nAgain, kCollideThreads), kCollideThreads, again, order, nAgain, ctx->contacts,
               ctx->bodies, ctx->motion, ctx->shapes, remaining, ctx->settings.speculative,
               0.005f * ctx->settings.length_unit, ctx->fresh, ctx->cnt);
        if (ctx->profiling) { HIP_TRY(hipEventRecord(ctx->evPool[ev + 1], ctx->stream)); }
        return refreshApply(ctx, again, nAgain, laned, impactLaned, nRest);
    }
    // Only the contacts collided again and their bodies go to the host; the callback reads no others.
    // Grown by half again whenever a longer list comes, so that steady stepping does not allocate.
    int e;
    if (nAgain > ctx->pinCap) {
        const uint32_t cap = std::max<uint32_t>(std::max<uint32_t>(nAgain, ctx->pinCap + ctx->pinCap / 2), 1024u);
        if ((e = growPinned(ctx, ctx->pinAgain, cap)) || (e = growPinned(ctx, ctx->pinContacts, cap)) ||
            (e = growPinned(ctx, ctx->pinBodies, 2 * (size_t)cap)) || (e = growPinned(ctx, ctx->pinFresh, cap))) {
            return e;
        }
        ctx->pinCap = cap;
    }
    const uint32_t waves = std::min<uint32_t>(nAgain, 4u * 8192u);
    LAUNCH(k_gather_again, (waves + 3) / 4, kThreads, again, nAgain, ctx->contacts, ctx->bodies, ctx->motion,
           ctx->pinAgain, ctx->pinContacts, ctx->pinBodies);
    HIP_TRY(hipEventRecord(ctx->ev[3], ctx->stream));
    // Every row but those of the contacts collided again is final now (rebased, new inverse masses):
    // they are packed while the host collides, and the others again after (by k_refresh_apply), as
    // prepareLanes packs all rows after a refresh.
    LAUNCH(k_pack, blocksFor(laned), kThreads, ctx->rowsList, laned, ctx->contacts, ctx->moving, ctx->mass, ctx->rows,
           ctx->rowOf);
    using Clock = std::chrono::steady_clock;
    const Clock::time_point t0 = Clock::now();
    HIP_TRY(hipEventSynchronize(ctx->ev[3]));
    const Clock::time_point t1 = Clock::now();
    if (cb(user, substep, remaining, nAgain, ctx->pinAgain, ctx->pinContacts, ctx->pinBodies, ctx->pinFresh) != 0) {
        return fail(ctx, TFS_E_CALLBACK, "the refresh callback failed at substep %d", substep);
    }
    if (ctx->profiling) {
        const Clock::time_point t2 = Clock::now();
        ctx->profLaunches[3] += 1;
        ctx->profMs[3] += std::chrono::duration<double, std::milli>(t1 - t0).count();
        ctx->profLaunches[4] += 1;
        ctx->profMs[4] += std::chrono::duration<double, std::milli>(t2 - t1).count();
        ctx->profBytes[4] += nAgain;
    }
    HIP_TRY(hipMemcpyAsync(ctx->fresh, ctx->pinFresh, sizeof(tfs_contact) * nAgain, hipMemcpyHostToDevice, ctx->stream));
    return refreshApply(ctx, again, nAgain, laned, impactLaned, nRest);
}

// After the narrowphase of a refresh: the new manifolds into the contacts, their rows packed again.
int refreshApply(tfs_ctx* ctx, const uint32_t* again, uint32_t nAgain, uint32_t laned, uint32_t impactLaned, uint32_t nRest) {
    const float matchSq = ctx->settings.speculative * ctx->settings.speculative;
    // A refreshed contact keeps its row; its points and axes are packed again in the same launch.
    LAUNCH(k_refresh_apply, blocksFor(nAgain), kThreads, again, nAgain, ctx->fresh,
           ctx->contacts, ctx->bodies, ctx->motion, ctx->mass, matchSq, ctx->ext, ctx->colors, ctx->moving, ctx->rows,
           ctx->rowOf);
    if (impactLaned > 0) {
        LAUNCH(k_impact_rows, blocksFor(impactLaned), kThreads, ctx->impactRowsList, impactLaned, ctx->rowOf,
               ctx->rows.hdr, ctx->impactHdr);
    }
    if (nRest > 0) { LAUNCH(k_rest_pack, blocksFor(nRest), kThreads, ctx->restEntries[0], nRest, ctx->contacts, ctx->mass, restPack(ctx)); }
    return 0;
}

// The solve block up to its substeps (src/world.cpp:1591-1657): step state and lists, prepare, impact
// list, colouring, row order and packing. Leaves the counts in `hc`.
int beginStep(tfs_ctx* ctx, float dt, DevCounts& hc) {
    const uint32_t nb = ctx->nb, nc = ctx->nc;
    const tfs_settings& s = ctx->settings;
    const Ctx c = makeCtx(s, dt);
    HIP_TRY(hipMemsetAsync(ctx->cnt, 0, sizeof(DevCounts), ctx->stream));
    HIP_TRY(hipMemsetAsync(ctx->bar, 0, sizeof(BarrierLines), ctx->stream));
    ctx->barrierBase = BarrierBase{};
    ctx->scanSlot = 0;

    // step state and lists
    const BodyScratch zero{ ctx->struck, ctx->deg, ctx->cursor, ctx->bdeg, ctx->bcursor, ctx->word };
    LAUNCH(k_step_init, blocksFor(nb), kThreads, ctx->bodies, nb, dt, ctx->motion, ctx->mass, ctx->moving, zero);
    LAUNCH(k_classify, blocksFor(nc), kThreads, ctx->contacts, nc, ctx->bodies, dt, s.speculative, ctx->fActive,
           ctx->fFast, ctx->fFastSolid, ctx->colors);
    int r;
    const ScanJob lists[3] = { { ctx->fActive, nullptr, nc, nullptr, ctx->active, &ctx->cnt->nActive },
                               { ctx->fFast, nullptr, nc, nullptr, ctx->fast, &ctx->cnt->nFast },
                               { ctx->fFastSolid, nullptr, nc, nullptr, ctx->fastSolid, &ctx->cnt->nFastSolid } };
    if ((r = scanJobs(ctx, lists, 3, nc))) { return r; }

    // prepare, impact
    LAUNCH(k_prepare, blocksFor(nc), kThreads, ctx->active, ctx->cnt, ctx->contacts, ctx->bodies, ctx->motion, ctx->mass,
           c.restitutionThreshold, ctx->struck, &ctx->cnt->pointsActive, &ctx->cnt->anyRestitution);
    LAUNCH(k_impact_flag, blocksFor(nc), kThreads, ctx->active, ctx->cnt, ctx->contacts, ctx->struck, ctx->impactFlag);
    if ((r = scan(ctx, ctx->impactFlag, &ctx->cnt->nActive, nc, nullptr, ctx->impactPos, &ctx->cnt->nImpact))) { return r; }
    LAUNCH(k_gather_list, blocksFor(nc), kThreads, ctx->impactPos, &ctx->cnt->nImpact, ctx->active, ctx->impact);
    LAUNCH(k_points_of, blocksFor(nc), kThreads, ctx->impact, &ctx->cnt->nImpact, ctx->contacts, &ctx->cnt->pointsImpact);

    // colouring
    LAUNCH(k_ep_count, blocksFor(nc), kThreads, ctx->active, ctx->cnt, ctx->contacts, ctx->moving, ctx->deg);
    if ((r = scan(ctx, ctx->deg, nullptr, nb, ctx->epStart, nullptr, &ctx->cnt->scanTotal))) { return r; }
    LAUNCH(k_ep_fill, blocksFor(nc), kThreads, ctx->active, ctx->cnt, ctx->contacts, ctx->moving, ctx->epStart,
           ctx->cursor, ctx->ep);
    LAUNCH(k_ep_rank, blocksFor(nb), kThreads, nb, ctx->deg, ctx->epStart, ctx->ep, ctx->rank);
    LAUNCH(k_color, blocksFor(nc, kColorThreads, 2048), kColorThreads, ctx->active, ctx->cnt, ctx->contacts, ctx->moving,
           ctx->rank, ctx->epStart, ctx->ep, ctx->word, ctx->colors, &ctx->cnt->ticket, ctx->cnt);

    // row order and packing
    if ((r = rowOrder(ctx, ctx->active, &ctx->cnt->nActive, nc, ctx->rowsList, ctx->cnt->colorStart,
                      ctx->cnt->colorPoints)) ||
        (r = rowOrder(ctx, ctx->impact, &ctx->cnt->nImpact, nc, ctx->impactRowsList, ctx->cnt->impactColorStart,
                      ctx->cnt->impactColorPoints))) {
        return r;
    }
    if ((r = syncCounts(ctx))) { return r; }
    hc = *ctx->hostCnt;
    return packRows(ctx, hc.colorStart[TFS_COLORS], hc.impactColorStart[TFS_COLORS], true);
}

void applyShape(tfs_ctx* ctx, const tfs_ctx::PassShape& x) {
    ctx->maxPassBlocks = std::max<uint32_t>(1, std::min(x.blocks, ctx->passBlockLimit));
    ctx->passTail = x.tailRows > 0;
    ctx->tailSegRows = x.tailRows > 0 ? x.tailRows : kTailSegRows;
    ctx->madeAxes = x.made;
    ctx->fused = x.fused;
}

// One substep in the fused launch (k_substep_fused): the same plans as the separate passes, one grid
// of at most fusedBlockLimit blocks for all four phases.
int substepFused(tfs_ctx* ctx, const Ctx& c, const DevCounts& hc) {
    PassPlan pb, pr;
    uint32_t unusedGrid = 0;
    double bytesB = 0.0, bytesR = 0.0;
    planPass(ctx, kBias, hc, true, hc.nImpact > 0, pb, unusedGrid, bytesB);
    planPass(ctx, kRelax, hc, true, hc.nImpact > 0, pr, unusedGrid, bytesR);
    const uint32_t grid = std::max<uint32_t>(1, std::min(ctx->maxPassBlocks, ctx->fusedBlockLimit));
    size_t ev = 0;
    if (ctx->profiling) {
        ev = ctx->pending.size() * 2;
        while (ctx->evPool.size() < ev + 2) {
            hipEvent_t x;
            HIP_TRY(hipEventCreate(&x));
            ctx->evPool.push_back(x);
        }
        ctx->pending.push_back({ 7, ev, rowBytes(kWarm, hc.colorStart[TFS_COLORS], hc.pointsActive) + bytesB + bytesR });
        HIP_TRY(hipEventRecord(ctx->evPool[ev], ctx->stream));
    }
    const BarrierBase base = ctx->barrierBase;
    const float4* mass = ctx->mass;
    const uint4* at = ctx->impactHdr;
    uint32_t nb = ctx->nb;
    void* args[] = { &ctx->bodies, &nb, &ctx->rows, (void*)&at, &pb, &pr, &ctx->motion, (void*)&mass, (void*)&c, &ctx->bdeg,
                     &ctx->bstart, &ctx->bep, &ctx->bar, (void*)&base, &ctx->cnt };
    const void* fn = ctx->madeAxes ? (const void*)k_substep_fused<true> : (const void*)k_substep_fused<false>;
    if (int e = launchGrid(ctx, fn, grid, args)) { return e; }
    advanceBarrier(ctx->barrierBase, grid, 3 + passBarriers(pb) + passBarriers(pr));
    if (ctx->profiling) { HIP_TRY(hipEventRecord(ctx->evPool[ev + 1], ctx->stream)); }
    return 0;
}

// The candidate a tuning step runs, and the bookkeeping after it.
uint32_t tuneCandidate(const tfs_ctx* ctx) {
    const uint32_t c = (uint32_t)ctx->tuneShapes.size();
    return (ctx->tuneStep + ctx->tuneStep / c) % c;
}

void tuneRecord(tfs_ctx* ctx, float ms) {
    const uint32_t c = (uint32_t)ctx->tuneShapes.size();
    ctx->tuneMs[tuneCandidate(ctx)].push_back(ms);
    if (++ctx->tuneStep < ctx->tuneRounds * c) { return; }
    uint32_t best = 0;
    float bestMs = 0.0f;
    for (uint32_t k = 0; k < c; ++k) {
        std::vector<float> t = ctx->tuneMs[k];
        std::sort(t.begin(), t.end());
        const float median = t[t.size() / 2];
        if (k == 0 || median < bestMs) {
            best = k;
            bestMs = median;
        }
    }
    applyShape(ctx, ctx->tuneShapes[best]);
    ctx->tuning = false;
}

// updateSleep (src/world.cpp:1174-1247) over the resident bodies and contacts: sleep timers and the
// kinematic wake, then hook-and-jump rounds until one links nothing - four rounds, then two that run
// only if the round before them linked something - then the island times and the sleep, which run only
// if the last round linked nothing: no host check in between. `it` receives the rounds launched; the
// last round's flag is copied to the host for finishSleep.
int launchSleep(tfs_ctx* ctx, float dt, int& it) {
    const uint32_t nb = ctx->nb, nc = ctx->nc;
    const tfs_settings& s = ctx->settings;
    it = 0;
    LAUNCH(k_sleep_time, blocksFor(nb), kThreads, ctx->bodies, nb, dt, s.sleep_threshold * s.length_unit, ctx->parent,
           ctx->islandTime);
    LAUNCH(k_sleep_kinematic, blocksFor(nc), kThreads, ctx->contacts, nc, ctx->bodies, ctx->edges);
    if (ctx->nJoints) {
        LAUNCH(k_sleep_joints, blocksFor(ctx->nJoints), kThreads, ctx->joints, ctx->nJoints, ctx->bodies, ctx->edges + nc);
    }
    const uint32_t ne = nc + ctx->nJoints;
    const bool worlds = ctx->nWorlds > 1;
    const uint32_t* bw = worlds ? ctx->bodyWorld : nullptr;
    const uint32_t* ws = worlds ? ctx->worldStart : nullptr;
    const int ahead = ctx->islandAhead;
    for (; it < ahead; ++it) {
        const uint32_t* prev = it >= 4 ? &ctx->cnt->changed[it - 1] : nullptr;
        LAUNCH(k_sleep_hook, blocksFor(ne), kThreads, ctx->edges, ne, ctx->parent, &ctx->cnt->changed[it], prev);
        LAUNCH(k_sleep_jump, blocksFor(nb), kThreads, ctx->parent, nb, prev);
    }
    const uint32_t* last = &ctx->cnt->changed[ahead - 1];
    LAUNCH(k_island_time, blocksFor(nb), kThreads, ctx->bodies, nb, ctx->parent, ctx->islandTime, bw, ws, last);
    LAUNCH(k_sleep_apply, blocksFor(nb), kThreads, ctx->bodies, nb, ctx->parent, ctx->islandTime, s.time_to_sleep,
           ctx->motion, ctx->mass, last);
    HIP_TRY(hipMemcpyAsync(&ctx->hostCnt->changed[ahead - 1], last, sizeof(uint32_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    return 0;
}

// The rare step whose islands need more rounds than launchSleep ran (seen once its flag reached the
// host): two rounds at a time, checking each pair, then the island times and the sleep.
int finishSleep(tfs_ctx* ctx, int it, bool timed) {
    const uint32_t nb = ctx->nb, ne = ctx->nc + ctx->nJoints;
    const tfs_settings& s = ctx->settings;
    const bool worlds = ctx->nWorlds > 1;
    const uint32_t* bw = worlds ? ctx->bodyWorld : nullptr;
    const uint32_t* ws = worlds ? ctx->worldStart : nullptr;
    const int ahead = ctx->islandAhead;
    if (ctx->hostCnt->changed[ahead - 1]) {
        for (;;) {
            if (it + 2 > kIslandRounds) {
                return fail(ctx, TFS_E_KERNEL, "the island pass did not settle in %d rounds", kIslandRounds);
            }
            for (int j = 0; j < 2; ++j, ++it) {
                LAUNCH(k_sleep_hook, blocksFor(ne), kThreads, ctx->edges, ne, ctx->parent, &ctx->cnt->changed[it], nullptr);
                LAUNCH(k_sleep_jump, blocksFor(nb), kThreads, ctx->parent, nb, nullptr);
            }
            HIP_TRY(hipMemcpyAsync(&ctx->hostCnt->changed[it - 1], &ctx->cnt->changed[it - 1], sizeof(uint32_t),
                                   hipMemcpyDeviceToHost, ctx->stream));
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            if (!ctx->hostCnt->changed[it - 1]) { break; }
        }
        LAUNCH(k_island_time, blocksFor(nb), kThreads, ctx->bodies, nb, ctx->parent, ctx->islandTime, bw, ws, nullptr);
        LAUNCH(k_sleep_apply, blocksFor(nb), kThreads, ctx->bodies, nb, ctx->parent, ctx->islandTime, s.time_to_sleep,
               ctx->motion, ctx->mass, nullptr);
        HIP_TRY(hipMemcpyAsync(&ctx->hostCnt->error, &ctx->cnt->error, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        if (timed) {
            HIP_TRY(hipEventRecord(ctx->ev[2], ctx->stream));
            HIP_TRY(hipEventSynchronize(ctx->ev[2]));
        } else {
            HIP_TRY(hipStreamSynchronize(ctx->stream));
        }
    }
    return 0;
}

int step(tfs_ctx* ctx, float dt, tfs_refresh_fn cb, void* user) {
    if (ctx->nShapes && ctx->nShapes != ctx->nb) {
        return fail(ctx, TFS_E_STATE, "%u body shapes for %u bodies", ctx->nShapes, ctx->nb);
    }
    if (ctx->nJoints && ctx->jointBodies > ctx->nb) {
        return fail(ctx, TFS_E_STATE, "the resident joints name body %u of %u: upload them again", ctx->jointBodies - 1,
                    ctx->nb);
    }
    if (ctx->nJoints && ctx->nWorlds > 1) {
        // each of the reference's worlds has a world body of its own; the joints here share one
        return fail(ctx, TFS_E_STATE, "joints with %u worlds side by side: step one world at a time", ctx->nWorlds);
    }
    if (ctx->tuning) { applyShape(ctx, ctx->tuneShapes[tuneCandidate(ctx)]); }
    const uint32_t nb = ctx->nb, nc = ctx->nc;
    const tfs_settings& s = ctx->settings;
    const Ctx c = makeCtx(s, dt);
    const int n = s.substeps > 1 ? s.substeps : 1;
    // a context without events (tfs_step_noalloc's) does not time its steps
    const bool timed = ctx->ev[0] != nullptr;
    if (timed) { HIP_TRY(hipEventRecord(ctx->ev[0], ctx->stream)); }
    DevCounts hc{};
    int r;
    if ((r = beginStep(ctx, dt, hc))) { return r; }
    const uint32_t laned = hc.colorStart[TFS_COLORS], impactLaned = hc.impactColorStart[TFS_COLORS];
    if (n > 1 && !cb && !ctx->nShapes) {
        bool needs = false;
        for (int i = 1; i < n; ++i) { needs = needs || (i == n / 2 ? hc.nFast : hc.nFastSolid) > 0; }
        if (needs) { return fail(ctx, TFS_E_CALLBACK, "fast contacts need a refresh callback"); }
    }
    ctx->restLevels[0] = ctx->restLevels[1] = ctx->restWide[0] = ctx->restWide[1] = 0;
    ctx->restBytes[0] = ctx->restBytes[1] = 0.0;
    if (hc.nActive > laned || hc.nImpact > impactLaned) {
        if ((r = buildRestLevels(ctx, 0, ctx->rowsList + laned, hc.nActive - laned)) ||
            (r = buildRestLevels(ctx, 1, ctx->impactRowsList + impactLaned, hc.nImpact - impactLaned))) {
            return r;
        }
        static_assert(offsetof(DevCounts, restNarrow) == offsetof(DevCounts, restLevels) + 2 * sizeof(uint32_t) &&
                          offsetof(DevCounts, restShape) == offsetof(DevCounts, restLevels) + 4 * sizeof(uint32_t),
                      "restLevels, restNarrow and restShape are read back together");
        HIP_TRY(hipMemcpyAsync(ctx->hostCnt->restLevels, ctx->cnt->restLevels, 12 * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipMemcpyAsync(&ctx->hostCnt->error, &ctx->cnt->error, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (ctx->hostCnt->error) { return kernelFail(ctx, ctx->hostCnt->error); }
        for (int w = 0; w < 2; ++w) {
            ctx->restLevels[w] = ctx->hostCnt->restLevels[w];
            const uint32_t narrow = ~ctx->hostCnt->restNarrow[w];  // ~0 -> no narrow level
            ctx->restWide[w] = std::min(narrow, ctx->restLevels[w]);
            const uint32_t* shape = ctx->hostCnt->restShape[w];
            ctx->restBytes[w] = restPassBytes(shape[0], shape[1], shape[2], shape[3]);
        }
        if (hc.nActive > laned) {
            const uint32_t nRest = hc.nActive - laned;
            LAUNCH(k_rest_pack, blocksFor(nRest), kThreads, ctx->restEntries[0], nRest, ctx->contacts, ctx->mass, restPack(ctx));
            // the impact rest is a part of the active rest: both lists take the active colouring
            // (src/world.cpp:1651-1657), so every impact rest contact has a place in the pack
            if (hc.nImpact > impactLaned) {
                LAUNCH(k_rest_impact_pos, blocksFor(hc.nImpact - impactLaned), kThreads, ctx->restEntries[1],
                       hc.nImpact - impactLaned, ctx->restPos);
            }
            LAUNCH(k_fill_u32, blocksFor(nb), kThreads, ctx->wrestDeg, nb, 0u);
            LAUNCH(k_fill_u32, blocksFor(nb), kThreads, ctx->wrestCursor, nb, 0u);
            LAUNCH(k_wrest_count, blocksFor(nRest), kThreads, ctx->restEntries[0], nRest, ctx->wrestDeg);
            if ((r = scan(ctx, ctx->wrestDeg, nullptr, nb, ctx->wrestStart, nullptr, &ctx->cnt->scanTotal))) { return r; }
            LAUNCH(k_wrest_fill, blocksFor(nRest), kThreads, ctx->restEntries[0], nRest, ctx->wrestStart, ctx->wrestCursor,
                   ctx->wrestEp);
            LAUNCH(k_bep_sort, blocksFor(nb), kThreads, nb, ctx->wrestDeg, ctx->wrestStart, ctx->wrestEp);
        }
    }

    // substeps: src/world.cpp:1658-1681. A substep with no rest list may run as one fused launch; with
    // no host callback to wait for, the whole loop may go out as one HIP graph.
    // With joints each substep's start is split around their warm start, and their solves go before the
    // contacts' in every pass: no fused substep and no graph.
    const bool joints = ctx->nJoints > 0;
    tfj::JCtx jc{ c.h, c.invH, makeSoft(minf(s.joint_hertz, 0.5f * c.invH), s.joint_damping_ratio, c.h) };
    if (joints) { LAUNCH(k_joints_begin, 1, 64, ctx->joints, ctx->nJoints, ctx->bodies, ctx->jointActive, ctx->cnt); }
    const bool fuse = !joints && ctx->fused && ctx->usePass && hc.nActive == laned && hc.nImpact == impactLaned;
    bool hostRefresh = false;
    for (int i = 1; i < n; ++i) { hostRefresh = hostRefresh || (!ctx->nShapes && (i == n / 2 ? hc.nFast : hc.nFastSolid) > 0); }
    const bool graph = !joints && ctx->useGraph && timed && !ctx->profiling && !ctx->tuning && !ctx->coopLaunch && !hostRefresh;
    if (n > 1 && ctx->nShapes && (hc.nFast > 0 || hc.nFastSolid > 0)) {
        LAUNCH(k_refresh_order, 2, 1024, ctx->fast, ctx->fastSolid, ctx->cnt, ctx->contacts, ctx->shapes, ctx->fastOrder,
               ctx->fastSolidOrder);
    }
    auto loop = [&]() -> int {
        for (int i = 0; i < n; ++i) {
            const bool half = i == n / 2;
            const uint32_t nAgain = half ? hc.nFast : hc.nFastSolid;
            int e;
            if (i > 0 && nAgain > 0) {
                if ((e = refresh(ctx, i, half, nAgain, (float)(n - i) * c.h, cb, user, hc))) { return e; }
            }
            if (fuse) {
                if ((e = substepFused(ctx, c, hc))) { return e; }
                continue;
            }
            if (joints) {
                if ((e = substepStart(ctx, c, laned, hc.pointsActive, hc.nActive > laned, 1))) { return e; }
                LAUNCH(k_joints_warm, 1, 64, ctx->joints, ctx->jointActive, ctx->cnt, ctx->bodies, ctx->motion,
                       ctx->worldMotion);
                if ((e = substepStart(ctx, c, laned, hc.pointsActive, hc.nActive > laned, 2))) { return e; }
                LAUNCH(k_joints_solve, 1, 64, ctx->joints, ctx->jointActive, ctx->cnt, ctx->bodies, ctx->motion,
                       ctx->worldMotion, jc, 1);
            } else if ((e = substepStart(ctx, c, laned, hc.pointsActive, hc.nActive > laned, 3))) {
                return e;
            }
            if ((e = solvePass(ctx, kBias, hc, laned, impactLaned, c))) { return e; }
            LAUNCH(k_integrate_positions, blocksFor(nb, kBodyThreads), kBodyThreads, ctx->bodies, nb, ctx->motion, c);
            if (joints) {
                LAUNCH(k_joints_solve, 1, 64, ctx->joints, ctx->jointActive, ctx->cnt, ctx->bodies, ctx->motion,
                       ctx->worldMotion, jc, 0);
            }
            if ((e = solvePass(ctx, kRelax, hc, laned, impactLaned, c))) { return e; }
        }
        return 0;
    };
    if (ctx->tuning) { HIP_TRY(hipEventRecord(ctx->ev[4], ctx->stream)); }
    if (graph) {
        HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
        r = loop();
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
        if (r || ec != hipSuccess) {
            if (g) { (void)hipGraphDestroy(g); }
            return r ? r : fail(ctx, TFS_E_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(ec));
        }
        // The last step's executable graph takes this step's parameters when its topology is the same
        // (the same launches in the same order); otherwise it is made again.
        bool updated = false;
        if (ctx->graphExec) {
            hipGraphExecUpdateResult how;
            hipGraphNode_t errNode = nullptr;
            updated = hipGraphExecUpdate(ctx->graphExec, g, &errNode, &how) == hipSuccess && how == hipGraphExecUpdateSuccess;
            if (!updated) {
                (void)hipGetLastError();
                HIP_TRY(hipGraphExecDestroy(ctx->graphExec));
                ctx->graphExec = nullptr;
            }
        }
        if (!updated) {
            const hipError_t ei = hipGraphInstantiate(&ctx->graphExec, g, nullptr, nullptr, 0);
            if (ei != hipSuccess) {
                (void)hipGraphDestroy(g);
                ctx->graphExec = nullptr;
                return fail(ctx, TFS_E_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(ei));
            }
        }
        (void)hipGraphDestroy(g);
        HIP_TRY(hipGraphLaunch(ctx->graphExec, ctx->stream));
    } else if ((r = loop())) {
        return r;
    }
    if (ctx->tuning) { HIP_TRY(hipEventRecord(ctx->ev[5], ctx->stream)); }
    // restituteBatches + applyRestitution (src/world.cpp:1682-1683): nothing to do unless some active
    // contact has restitution.
    if (hc.anyRestitution && ((r = runRows(ctx, kRestitute, nullptr, hc.colorStart, hc.colorPoints, c)) ||
                              (r = runRest(ctx, kRestitute, ctx->rowsList, laned, hc.nActive, c)))) {
        return r;
    }
    LAUNCH(k_limit_speeds, blocksFor(nb), kThreads, ctx->bodies, nb, ctx->motion, c);
    if (joints) {
        LAUNCH(k_joints_end, 1, 64, ctx->joints, ctx->jointActive, ctx->cnt, ctx->bodies, ctx->motion, ctx->worldMotion, dt);
        LAUNCH(k_joints_mass, blocksFor(nb), kThreads, ctx->bodies, nb, ctx->cnt);
    }
    LAUNCH(k_unpack, blocksFor(laned), kThreads, laned, ctx->rows, ctx->contacts);
    if (hc.nActive > laned) {
        LAUNCH(k_rest_unpack, blocksFor(hc.nActive - laned), kThreads, ctx->restEntries[0], hc.nActive - laned,
               restPack(ctx), ctx->contacts);
    }
    LAUNCH(k_writeback, blocksFor(nb), kThreads, ctx->bodies, nb, ctx->motion);
    if (timed) { HIP_TRY(hipEventRecord(ctx->ev[1], ctx->stream)); }

    // reportTouches (src/world.cpp:1693, 1424-1468): on the contacts and poses the solve block left, before
    // the sleep pass changes which bodies are awake
    if (nc > 0) {
        LAUNCH(k_touches, blocksFor(nc), kThreads, ctx->contacts, nc, ctx->bodies, s.hit_threshold * s.length_unit,
               0.25f * s.speculative, ctx->touchFlag, ctx->hitFlag, ctx->hitData, &ctx->cnt->touching);
        const ScanJob touch[2] = { { ctx->touchFlag, nullptr, nc, nullptr, ctx->touchList, &ctx->cnt->nTouchEvents },
                                   { ctx->hitFlag, nullptr, nc, nullptr, ctx->hitList, &ctx->cnt->nHits } };
        if ((r = scanJobs(ctx, touch, 2, nc))) { return r; }
        LAUNCH(k_gather_hits, blocksFor(nc), kThreads, ctx->hitList, &ctx->cnt->nHits, ctx->hitData, ctx->hits);
        LAUNCH(k_touch_events, blocksFor(nc), kThreads, ctx->touchList, &ctx->cnt->nTouchEvents, ctx->contacts);
    }

    // sleep and islands (unless the caller runs them later, tfs_set_step_sleep)
    int it = 0;
    if (ctx->stepSleep && (r = launchSleep(ctx, dt, it))) { return r; }
    HIP_TRY(hipMemcpyAsync(&ctx->hostCnt->error, &ctx->cnt->error, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    static_assert(offsetof(DevCounts, nHits) == offsetof(DevCounts, nTouchEvents) + sizeof(uint32_t) &&
                      offsetof(DevCounts, touching) == offsetof(DevCounts, nTouchEvents) + 2 * sizeof(uint32_t),
                  "the touch counts are read back together");
    HIP_TRY(hipMemcpyAsync(&ctx->hostCnt->nTouchEvents, &ctx->cnt->nTouchEvents, 3 * sizeof(uint32_t),
                           hipMemcpyDeviceToHost, ctx->stream));
    if (timed) {
        HIP_TRY(hipEventRecord(ctx->ev[2], ctx->stream));
        HIP_TRY(hipEventSynchronize(ctx->ev[2]));
    } else {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
    }
    ctx->touchCounts[0] = ctx->hostCnt->nTouchEvents;
    ctx->touchCounts[1] = ctx->hostCnt->nHits;
    ctx->touchCounts[2] = ctx->hostCnt->touching;
    if (ctx->stepSleep && (r = finishSleep(ctx, it, timed))) { return r; }
    if (ctx->hostCnt->error) { return kernelFail(ctx, ctx->hostCnt->error); }
    if (timed) {
        HIP_TRY(hipEventElapsedTime(&ctx->solveMs, ctx->ev[0], ctx->ev[1]));
        HIP_TRY(hipEventElapsedTime(&ctx->sleepMs, ctx->ev[1], ctx->ev[2]));
    }
    if ((r = collectProfile(ctx))) { return r; }
    if (ctx->tuning) {
        // the candidate's time: the whole substep loop (refreshes, substep starts, passes, positions), per
        // contact point, so that steps of a pile still filling up compare fairly
        HIP_TRY(hipEventElapsedTime(&ctx->tunePassMs, ctx->ev[4], ctx->ev[5]));
        const uint32_t points = hc.pointsActive + hc.pointsImpact;
        if (hc.nActive > 0 && points > 0) { tuneRecord(ctx, ctx->tunePassMs / (float)points); }
        ctx->tunePassMs = 0.0f;
    }
    uint32_t nColors = 0;
    for (int col = 0; col < TFS_COLORS; ++col) {
        if (hc.colorStart[col + 1] > hc.colorStart[col]) { nColors = (uint32_t)col + 1; }
    }
    const uint32_t counts[8] = { hc.nActive, hc.pointsActive, hc.nImpact, hc.pointsImpact, nColors, hc.nActive - laned,
                                  ctx->restLevels[0], ctx->restLevels[1] };
    memcpy(ctx->lastCounts, counts, sizeof counts);
    return 0;
}

// ---- the fixed-contact model and its adjoint ----------------------------------------------------

void freeFixed(tfs_ctx* ctx) {
    for (void* p : { (void*)ctx->ckBody, (void*)ctx->ckRow, (void*)ctx->tapeBuf, (void*)ctx->adj.bar,
                     (void*)ctx->adj.dpBar, (void*)ctx->adj.comBar, (void*)ctx->adj.impBar, (void*)ctx->adj.sepBar,
                     (void*)ctx->adj.muBar, (void*)ctx->dlcom, (void*)ctx->muContact }) {
        if (p) { (void)hipFree(p); }
    }
    ctx->ckBody = ctx->ckRow = ctx->tapeBuf = nullptr;
    ctx->adj = Adjoint{};
    ctx->dlcom = ctx->muContact = nullptr;
    ctx->fixedOn = false;
    ctx->fixedSteps = ctx->fixedMax = ctx->fixedLeft = 0;
}

// The tape of pass `i` of a step (pass 2s is substep s's biased pass, 2s + 1 its relax pass).
Tape tapeOf(const tfs_ctx* ctx, int i) {
    const size_t cap = ctx->rows.cap;
    float4* base = ctx->tapeBuf + (size_t)i * 9 * cap;
    return Tape{ base, base + 4 * cap, base + 8 * cap, (uint32_t)cap };
}

Checkpoint checkpointOf(const tfs_ctx* ctx, uint32_t k) {
    return Checkpoint{ ctx->ckBody + (size_t)k * 3 * ctx->nb, ctx->ckRow + (size_t)k * 5 * ctx->rows.cap };
}

// One step of the fixed-contact model, recording each pass into the tape when `record`.
int fixedStep(tfs_ctx* ctx, bool record) {
    const uint32_t nb = ctx->nb;
    const DevCounts& hc = ctx->fixedCounts;
    const uint32_t laned = hc.colorStart[TFS_COLORS];
    const Ctx c = makeCtx(ctx->settings, ctx->fixedDt);
    const int n = ctx->settings.substeps > 1 ? ctx->settings.substeps : 1;
    int r;
    auto pass = [&](int mode, int slot) {
        const Tape t = record ? tapeOf(ctx, slot) : Tape{};
        return ctx->usePass ? runPass(ctx, mode, hc, true, false, c, t)
                            : runRows(ctx, mode, nullptr, hc.colorStart, hc.colorPoints, c, t);
    };
    for (int i = 0; i < n; ++i) {
        if ((r = substepStart(ctx, c, laned, hc.pointsActive, false, 3)) || (r = pass(kBias, 2 * i))) { return r; }
        LAUNCH(k_fx_positions, blocksFor(nb), kThreads, ctx->bodies, nb, ctx->motion, c.h);
        if ((r = pass(kRelax, 2 * i + 1))) { return r; }
    }
    LAUNCH(k_fx_rebase, blocksFor(laned), kThreads, ctx->rows, laned, ctx->motion);
    LAUNCH(k_fx_reset_dp, blocksFor(nb), kThreads, nb, ctx->motion);
    LAUNCH(k_writeback, blocksFor(nb), kThreads, ctx->bodies, nb, ctx->motion);
    return 0;
}

// The reverse of one pass: the colours last to first, in one launch with a grid barrier (or one launch per
// colour where the pass launch is off).
int passAdjoint(tfs_ctx* ctx, int mode, int slot, const Ctx& c) {
    const DevCounts& hc = ctx->fixedCounts;
    const Tape tape = tapeOf(ctx, slot);
    const float4* mass = ctx->mass;
    const bool biased = mode == kBias;
    PassPlan plan{};
    uint32_t widest = 0;
    for (int col = TFS_COLORS - 1; col >= 0; --col) {
        if (hc.colorStart[col + 1] <= hc.colorStart[col]) { continue; }
        plan.begin[plan.n] = hc.colorStart[col];
        plan.end[plan.n] = hc.colorStart[col + 1];
        widest = std::max(widest, plan.end[plan.n] - plan.begin[plan.n]);
        ++plan.n;
    }
    if (plan.n == 0) { return 0; }
    size_t ev = 0;
    if (ctx->profiling) {
        // bytes per row of the reverse sweep: the solve's row (224 + 296n), its record (64 + 16 + 16n), the
        // two bodies' velocity adjoints read and written (128), per point the impulse adjoints read and
        // written (32) and the separation adjoint (4), the friction adjoint (8)
        uint32_t rowsIn = 0, pointsIn = 0;
        for (int col = 0; col < TFS_COLORS; ++col) {
            rowsIn += hc.colorStart[col + 1] - hc.colorStart[col];
            pointsIn += hc.colorPoints[col];
        }
        ev = ctx->pending.size() * 2;
        while (ctx->evPool.size() < ev + 2) {
            hipEvent_t x;
            HIP_TRY(hipEventCreate(&x));
            ctx->evPool.push_back(x);
        }
        ctx->pending.push_back({ 6, ev, 440.0 * rowsIn + 348.0 * pointsIn });
        HIP_TRY(hipEventRecord(ctx->evPool[ev], ctx->stream));
    }
    const void* fn = ctx->madeAxes ? (biased ? (const void*)k_pass_adj<true, true> : (const void*)k_pass_adj<false, true>)
                                   : (biased ? (const void*)k_pass_adj<true, false> : (const void*)k_pass_adj<false, false>);
    if (ctx->usePass) {
        const uint32_t grid = std::max<uint32_t>(1, std::min<uint32_t>((widest + 63) / 64, ctx->maxPassBlocks));
        const BarrierBase base = ctx->barrierBase;
        void* args[] = { &ctx->rows, &plan, (void*)&tape, (void*)&mass, (void*)&c, &ctx->adj, &ctx->bar, (void*)&base,
                         &ctx->cnt };
        if (int e = launchGrid(ctx, fn, grid, args)) { return e; }
        advanceBarrier(ctx->barrierBase, grid, plan.n - 1);
        if (ctx->profiling) { HIP_TRY(hipEventRecord(ctx->evPool[ev + 1], ctx->stream)); }
        return 0;
    }
    for (uint32_t k = 0; k < plan.n; ++k) {
        PassPlan one{};
        one.n = 1;
        one.begin[0] = plan.begin[k];
        one.end[0] = plan.end[k];
        const uint32_t grid = (one.end[0] - one.begin[0] + kPassThreads - 1) / kPassThreads;
        const BarrierBase base = ctx->barrierBase;
        void* args[] = { &ctx->rows, &one, (void*)&tape, (void*)&mass, (void*)&c, &ctx->adj, &ctx->bar, (void*)&base,
                         &ctx->cnt };
        HIP_TRY(hipLaunchKernel(fn, dim3(grid), dim3(kPassThreads), args, 0, ctx->stream));
    }
    if (ctx->profiling) { HIP_TRY(hipEventRecord(ctx->evPool[ev + 1], ctx->stream)); }
    return 0;
}

int checkErrors(tfs_ctx* ctx) {
    HIP_TRY(hipMemcpyAsync(&ctx->hostCnt->error, &ctx->cnt->error, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->hostCnt->error) { return kernelFail(ctx, ctx->hostCnt->error); }
    return 0;
}

int fixedBegin(tfs_ctx* ctx, float dt, uint32_t maxSteps) {
    freeFixed(ctx);
    DevCounts hc{};
    int r;
    if ((r = beginStep(ctx, dt, hc))) { return r; }
    const uint32_t laned = hc.colorStart[TFS_COLORS];
    if (hc.nActive > laned) {
        return fail(ctx, TFS_E_STATE, "the fixed-contact model needs every active contact in a colour (%u of %u are not)",
                    hc.nActive - laned, hc.nActive);
    }
    const size_t cap = ctx->rows.cap, nb = ctx->nb, passes = 2 * (size_t)std::max(ctx->settings.substeps, 1);
    auto grab = [&](void** p, size_t bytes) { return hipMalloc(p, std::max<size_t>(bytes, 16)) == hipSuccess; };
    if (!grab((void**)&ctx->ckBody, maxSteps * 3 * nb * sizeof(float4)) ||
        !grab((void**)&ctx->ckRow, maxSteps * 5 * cap * sizeof(float4)) ||
        !grab((void**)&ctx->tapeBuf, passes * 9 * cap * sizeof(float4)) ||
        !grab((void**)&ctx->adj.bar, 2 * nb * sizeof(float4)) || !grab((void**)&ctx->adj.dpBar, nb * sizeof(float4)) ||
        !grab((void**)&ctx->adj.comBar, nb * sizeof(float4)) ||
        !grab((void**)&ctx->adj.impBar, 4 * cap * sizeof(float4)) ||
        !grab((void**)&ctx->adj.sepBar, 4 * cap * sizeof(float)) || !grab((void**)&ctx->adj.muBar, cap * sizeof(float)) ||
        !grab((void**)&ctx->dlcom, 3 * nb * sizeof(float)) ||
        !grab((void**)&ctx->muContact, std::max<size_t>(ctx->nc, 1) * sizeof(float))) {
        freeFixed(ctx);
        return fail(ctx, TFS_E_DEVICE, "out of device memory for %u checkpoints", maxSteps);
    }
    uint32_t nColors = 0;
    for (int col = 0; col < TFS_COLORS; ++col) {
        if (hc.colorStart[col + 1] > hc.colorStart[col]) { nColors = (uint32_t)col + 1; }
    }
    const uint32_t counts[8] = { hc.nActive, hc.pointsActive, 0, 0, nColors, 0, 0, 0 };
    memcpy(ctx->lastCounts, counts, sizeof counts);
    ctx->fixedCounts = hc;
    ctx->fixedDt = dt;
    ctx->fixedMax = maxSteps;
    ctx->fixedOn = true;
    return checkErrors(ctx);
}

}  // namespace

// ---- the step that allocates nothing (tfs_step_noalloc) ------------------------------------------
//
// Its scratch: every device buffer a step of up to kNoAllocBodies bodies and kNoAllocContacts contacts
// needs (the body and contact arrays the step works on, rows, lists, colouring and island scratch,
// scan status words, counters, the grid barrier's lines), carved once per device out of this array.
// Host code never names the array itself: it uses the address hipGetSymbolAddress gives for the
// current device.
constexpr uint32_t kNoAllocBodies = TFS_NOALLOC_MAX_BODIES;
constexpr uint32_t kNoAllocContacts = TFS_NOALLOC_MAX_CONTACTS;
constexpr size_t kNoAllocScratchBytes = (size_t)28 << 20;
constexpr int kNoAllocDevices = 16;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Rejects what would send a kernel out of bounds: a contact naming a body out of range or holding
// more than 4 points, a shape that is not a sphere, box or half-space of finite positive size.
__global__ void k_check_input(const tfs_contact* contacts, uint32_t nc, uint32_t nb, const tfs_shape* shapes,
                              uint32_t ns, DevCounts* cnt) {
    for (uint32_t i = gid(); i < nc || i < ns; i += gstride()) {
        bool bad = false;
        if (i < nc) {
            const tfs_contact& c = contacts[i];
            bad = c.body_a >= nb || c.body_b >= nb || c.count < 0 || c.count > 4;
        }
        if (i < ns) {
            const tfs_shape& x = shapes[i];
            const bool ok = x.kind == TFS_SHAPE_HALF_SPACE ||
                            (x.kind == TFS_SHAPE_SPHERE && x.radius > 0.0f && x.radius <= 3.4e38f) ||
                            (x.kind == TFS_SHAPE_BOX && x.half[0] > 0.0f && x.half[0] <= 3.4e38f && x.half[1] > 0.0f &&
                             x.half[1] <= 3.4e38f && x.half[2] > 0.0f && x.half[2] <= 3.4e38f);
            bad = bad || !ok;
        }
        if (bad) { atomicOr(&cnt->error, 32u); }
    }
}

namespace {

// One context per device for tfs_step_noalloc, set up on its first call on that device, its device
// memory all carved from that device's copy of `scratch`; counts come back through plain host memory.
tfs_ctx g_noalloc[kNoAllocDevices];
DevCounts g_noallocHostCnt[kNoAllocDevices];
bool g_noallocReady[kNoAllocDevices];

int noallocSetUp(int device) {
    tfs_ctx* ctx = &g_noalloc[device];
    ctx->error.reserve(512);   // error messages then fit without allocating
    tfs_default_settings(&ctx->settings);
    ctx->device = device;
    ctx->stream = nullptr;
    void* base = nullptr;
    HIP_TRY(hipGetSymbolAddress(&base, HIP_SYMBOL(scratch)));
    if (!base) { return fail(ctx, TFS_E_DEVICE, "no address for the scratch array on device %d", device); }
    ctx->arena = (char*)base;
    ctx->arenaCap = kNoAllocScratchBytes;
    ctx->arenaUsed = 0;
    int r;
    if ((r = carve(ctx, (void**)&ctx->cnt, sizeof(DevCounts))) || (r = carve(ctx, (void**)&ctx->bar, sizeof(BarrierLines)))) {
        return r;
    }
    ctx->hostCnt = &g_noallocHostCnt[device];
    int perCu = 0, cus = 0, coop = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, (const void*)k_pass<kRelax, false, true>, kPassThreads, 0));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    HIP_TRY(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device));
    ctx->passBlockLimit = (uint32_t)std::max(1, perCu * cus / 2);
    ctx->maxPassBlocks = std::min<uint32_t>(ctx->passBlockLimit, 128u);
    ctx->usePass = coop != 0 && perCu > 0;
    if ((r = allocBodies(ctx, kNoAllocBodies)) || (r = allocContacts(ctx, kNoAllocContacts))) { return r; }
    HIP_TRY(hipStreamSynchronize(nullptr));
    g_noallocReady[device] = true;
    return 0;
}

}  // namespace

// tfs_pose of each body, for tfs_download_poses.
__global__ void k_poses(const tfs_body* bodies, uint32_t n, tfs_pose* out) {
    for (uint32_t i = gid(); i < n; i += gstride()) {
        const tfs_body& b = bodies[i];
        tfs_pose x;
        memset(&x, 0, sizeof x);
        for (int k = 0; k < 3; ++k) {
            x.p[k] = b.p[k];
            x.com[k] = b.com[k];
            x.v[k] = b.v[k];
            x.w[k] = b.w[k];
        }
        for (int k = 0; k < 4; ++k) { x.q[k] = b.q[k]; }
        x.type = b.type;
        x.awake = b.awake;
        out[i] = x;
    }
}

// The solves' max and min (maxq, minq) over given pairs, for tfs_check_minmax.
__global__ void k_minmax(const float* a, const float* b, uint32_t n, float* out) {
    for (uint32_t i = gid(); i < n; i += gstride()) {
        out[2 * i] = maxq(a[i], b[i]);
        out[2 * i + 1] = minq(a[i], b[i]);
    }
}

extern "C" {

int tfs_abi_version(void) { return TFS_ABI_VERSION; }

int tfs_check_minmax(int device, const float* a, const float* b, uint32_t n, float* out) {
    if (!a || !b || !out) { return TFS_E_ARG; }
    if (n == 0) { return 0; }
    if (hipSetDevice(device) != hipSuccess) { return TFS_E_DEVICE; }
    float *da = nullptr, *db = nullptr, *dout = nullptr;
    int r = 0;
    if (hipMalloc((void**)&da, sizeof(float) * n) != hipSuccess || hipMalloc((void**)&db, sizeof(float) * n) != hipSuccess ||
        hipMalloc((void**)&dout, 2 * sizeof(float) * n) != hipSuccess ||
        hipMemcpy(da, a, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(db, b, sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess) {
        r = TFS_E_DEVICE;
    }
    if (!r) {
        hipLaunchKernelGGL(k_minmax, dim3(std::min<uint32_t>((n + kThreads - 1) / kThreads, 4096u)), dim3(kThreads), 0, 0,
                           da, db, n, dout);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpy(out, dout, 2 * sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess) {
            r = TFS_E_DEVICE;
        }
    }
    for (float* p : { da, db, dout }) {
        if (p) { (void)hipFree(p); }
    }
    return r;
}
size_t tfs_sizeof_body(void) { return sizeof(tfs_body); }
size_t tfs_sizeof_contact(void) { return sizeof(tfs_contact); }
size_t tfs_sizeof_settings(void) { return sizeof(tfs_settings); }
size_t tfs_sizeof_joint(void) { return sizeof(tfs_joint); }

void tfs_default_settings(tfs_settings* s) {
    if (!s) { return; }
    *s = tfs_settings{};
    s->gravity[1] = -9.81f;
    s->length_unit = 1.0f;
    s->substeps = 4;
    s->contact_hertz = 30.0f;
    s->contact_damping_ratio = 10.0f;
    s->max_pushout_speed = 3.0f;
    s->max_linear_speed = 1000.0f;
    s->max_angular_speed = 1000.0f;
    s->restitution_threshold = 1.0f;
    s->sleep_threshold = 0.05f;
    s->time_to_sleep = 0.5f;
    s->speculative = 0.02f;
    s->hit_threshold = 1.0f;
    s->contact_reuse_distance = 0.0005f;
    s->contact_reuse_angle = 0.035f;
    s->joint_hertz = 60.0f;
    s->joint_damping_ratio = 2.0f;
}

int tfs_create(int device, uint32_t maxBodies, uint32_t maxContacts, tfs_ctx** out) {
    if (!out) { return TFS_E_ARG; }
    *out = nullptr;
    tfs_ctx* ctx = new tfs_ctx();
    tfs_default_settings(&ctx->settings);
    ctx->device = device;
    int r = 0;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return TFS_E_DEVICE;
    }
    for (hipEvent_t& e : ctx->ev) {
        if (hipEventCreate(&e) != hipSuccess) { r = TFS_E_DEVICE; }
    }
    if (!r && hipMalloc((void**)&ctx->cnt, sizeof(DevCounts)) != hipSuccess) { r = TFS_E_DEVICE; }
    if (!r && hipMalloc((void**)&ctx->bar, sizeof(BarrierLines)) != hipSuccess) { r = TFS_E_DEVICE; }
    if (!r && hipHostMalloc((void**)&ctx->hostCnt, sizeof(DevCounts), hipHostMallocDefault) != hipSuccess) { r = TFS_E_DEVICE; }
    if (!r) {
        int perCu = 0, cus = 0, coop = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, (const void*)k_pass<kRelax, false, true>, kPassThreads, 0) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
            hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device) != hipSuccess) {
            r = TFS_E_DEVICE;
        }
        // Every block of a grid-barrier pass must be resident at once (half the occupancy limit keeps well
        // inside that). The barrier's cost grows with the block count (about a microsecond at 64
        // blocks, 4 at 256), so the pass uses at most 128 blocks, one row per thread up to 32k rows
        // a colour.
        ctx->passBlockLimit = (uint32_t)std::max(1, perCu * cus / 2);
        ctx->maxPassBlocks = std::min<uint32_t>(ctx->passBlockLimit, 128u);
        ctx->usePass = coop != 0 && perCu > 0;
        int perCuFused = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCuFused, (const void*)k_substep_fused<false>, kPassThreads, 0) !=
            hipSuccess) {
            r = TFS_E_DEVICE;
        }
        ctx->fusedBlockLimit = (uint32_t)std::max(0, perCuFused * cus / 2);
        // Tuning knobs, read once: TWOFOLD_PASS=launch|coop, TWOFOLD_PASS_BLOCKS=n (at most the resident
        // limit of a grid-barrier launch), TWOFOLD_AXES=packed|made, TWOFOLD_TAIL=on|off,
        // TWOFOLD_ISLAND_AHEAD=n (1..6).
        if (const char* e = getenv("TWOFOLD_PASS")) { ctx->usePass = ctx->usePass && strcmp(e, "launch") != 0; }
        if (const char* e = getenv("TWOFOLD_PASS_BLOCKS")) {
            ctx->maxPassBlocks = (uint32_t)std::max(1, std::min(atoi(e), (int)ctx->passBlockLimit));
        }
        if (const char* e = getenv("TWOFOLD_ISLAND_AHEAD")) {
            ctx->islandAhead = std::max(1, std::min(atoi(e), kIslandAhead));
        }
        if (const char* e = getenv("TWOFOLD_FUSED")) { ctx->fused = strcmp(e, "on") == 0 && ctx->fusedBlockLimit > 0; }
        if (const char* e = getenv("TWOFOLD_GRAPH")) { ctx->useGraph = strcmp(e, "on") == 0; }
        if (const char* e = getenv("TWOFOLD_AXES")) { ctx->madeAxes = strcmp(e, "made") == 0; }
        if (const char* e = getenv("TWOFOLD_TAIL")) { ctx->passTail = strcmp(e, "off") != 0; }
        if (const char* e = getenv("TWOFOLD_COOP")) { ctx->coopLaunch = strcmp(e, "off") != 0; }
    }
    if (!r) { r = allocBodies(ctx, maxBodies ? maxBodies : 1); }
    if (!r) { r = allocContacts(ctx, maxContacts ? maxContacts : 1); }
    if (r) {
        tfs_destroy(ctx);
        return r;
    }
    *out = ctx;
    return 0;
}

void tfs_destroy(tfs_ctx* ctx) {
    if (!ctx) { return; }
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) { (void)hipStreamSynchronize(ctx->stream); }
    void* ptrs[] = { ctx->bodies, ctx->motion, ctx->mass, ctx->moving, ctx->struck, ctx->deg, ctx->epStart, ctx->cursor,
                     ctx->word, ctx->parent, ctx->islandTime, ctx->contacts, ctx->fresh, ctx->shapes, ctx->fActive, ctx->fFast,
                     ctx->fFastSolid, ctx->active, ctx->fast, ctx->fastSolid, ctx->impactFlag, ctx->impactPos,
                     ctx->impact, ctx->rowsList, ctx->impactRowsList, ctx->colors, ctx->ep, ctx->rank, ctx->rowOf,
                     ctx->hist, ctx->status, ctx->cnt, ctx->bar, ctx->rows.hdr, ctx->rows.nrm,
                     ctx->rows.tg1, ctx->rows.tg2, ctx->rows.pa, ctx->rows.pb, ctx->rows.imp, ctx->rows.peak,
                     ctx->rows.ax, ctx->rows.sep, ctx->impactHdr, ctx->bdeg, ctx->bstart, ctx->bcursor, ctx->bep,
                     ctx->edges, ctx->bodyWorld, ctx->refreshing, ctx->worldStart, ctx->restTracked,
                     ctx->restDeg, ctx->restStart, ctx->restCursor, ctx->restEp, ctx->restRank, ctx->restLevel,
                     ctx->restCount, ctx->restWord, ctx->restEntries[0], ctx->restEntries[1], ctx->restLevelStart[0],
                     ctx->restLevelStart[1], ctx->restH, ctx->restP, ctx->restPeak, ctx->restAx, ctx->restPre, ctx->restPos, ctx->wrestDeg,
                     ctx->wrestStart, ctx->wrestCursor, ctx->wrestEp, ctx->touchFlag, ctx->hitFlag, ctx->touchList,
                     ctx->hitList, ctx->hitData, ctx->hits, ctx->ext, ctx->contactsNext, ctx->extNext, ctx->pairs,
                     ctx->pairLo, ctx->pairHi, ctx->pairKeep, ctx->pairStart, ctx->manifolds, ctx->ckBody, ctx->ckRow, ctx->tapeBuf, ctx->adj.bar, ctx->adj.dpBar,
                     ctx->adj.comBar, ctx->adj.impBar, ctx->adj.sepBar, ctx->adj.muBar, ctx->dlcom, ctx->muContact,
                     ctx->fastOrder, ctx->fastSolidOrder, ctx->joints, ctx->jointActive, ctx->worldMotion };
    for (void* p : ptrs) {
        if (p) { (void)hipFree(p); }
    }
    if (ctx->hostCnt) { (void)hipHostFree(ctx->hostCnt); }
    for (void* h : { (void*)ctx->pinBodies, (void*)ctx->pinContacts, (void*)ctx->pinAgain, (void*)ctx->pinFresh }) {
        if (h) { (void)hipHostFree(h); }
    }
    for (hipEvent_t e : ctx->ev) {
        if (e) { (void)hipEventDestroy(e); }
    }
    for (hipEvent_t e : ctx->evPool) { (void)hipEventDestroy(e); }
    if (ctx->graphExec) { (void)hipGraphExecDestroy(ctx->graphExec); }
    if (ctx->stream) { (void)hipStreamDestroy(ctx->stream); }
    delete ctx;
}

const char* tfs_last_error(const tfs_ctx* ctx) { return ctx ? ctx->error.c_str() : "no context"; }

int tfs_set_settings(tfs_ctx* ctx, const tfs_settings* s) {
    if (!ctx || !s) { return fail(ctx, TFS_E_ARG, "null argument"); }
    if (s->substeps < 1 || s->substeps > 64) { return fail(ctx, TFS_E_ARG, "substeps %d out of range", s->substeps); }
    ctx->settings = *s;
    return 0;
}

int tfs_upload_bodies(tfs_ctx* ctx, const tfs_body* bodies, uint32_t n) {
    if (!ctx || (!bodies && n)) { return fail(ctx, TFS_E_ARG, "null argument"); }
    ctx->fixedOn = false;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n > ctx->capB) {
        int r = allocBodies(ctx, n);
        if (r) { return r; }
    }
    if (n) { HIP_TRY(hipMemcpy(ctx->bodies, bodies, sizeof(tfs_body) * n, hipMemcpyHostToDevice)); }
    ctx->nb = n;
    ctx->haveBodies = true;
    ctx->nWorlds = 1;
    ctx->hostBodyWorld.clear();
    return 0;
}

int tfs_set_worlds(tfs_ctx* ctx, uint32_t n_worlds, const uint32_t* body_start) {
    if (!ctx || !body_start || n_worlds == 0) { return fail(ctx, TFS_E_ARG, "null argument or no world"); }
    ctx->fixedOn = false;
    if (!ctx->haveBodies) { return fail(ctx, TFS_E_STATE, "upload bodies before setting worlds"); }
    if (body_start[0] != 0 || body_start[n_worlds] != ctx->nb) {
        return fail(ctx, TFS_E_ARG, "world ranges must cover the %u bodies", ctx->nb);
    }
    std::vector<uint32_t> map(ctx->nb);
    for (uint32_t w = 0; w < n_worlds; ++w) {
        if (body_start[w + 1] < body_start[w]) { return fail(ctx, TFS_E_ARG, "world %u has a negative range", w); }
        for (uint32_t i = body_start[w]; i < body_start[w + 1]; ++i) { map[i] = w; }
    }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n_worlds > ctx->worldCap) {
        int r;
        if ((r = grow(ctx, ctx->refreshing, n_worlds)) || (r = grow(ctx, ctx->worldStart, (size_t)n_worlds + 1))) { return r; }
        ctx->worldCap = n_worlds;
    }
    if (ctx->nb) { HIP_TRY(hipMemcpy(ctx->bodyWorld, map.data(), sizeof(uint32_t) * ctx->nb, hipMemcpyHostToDevice)); }
    HIP_TRY(hipMemcpy(ctx->worldStart, body_start, sizeof(uint32_t) * ((size_t)n_worlds + 1), hipMemcpyHostToDevice));
    ctx->nWorlds = n_worlds;
    ctx->hostBodyWorld.swap(map);
    // Contacts are uploaded again after this, so that each is checked to stay inside one world.
    ctx->nc = 0;
    ctx->haveContacts = false;
    return 0;
}

int tfs_upload_contacts(tfs_ctx* ctx, const tfs_contact* contacts, uint32_t n) {
    if (!ctx || (!contacts && n)) { return fail(ctx, TFS_E_ARG, "null argument"); }
    ctx->fixedOn = false;
    if (!ctx->haveBodies) { return fail(ctx, TFS_E_STATE, "upload bodies before contacts"); }
    for (uint32_t i = 0; i < n; ++i) {
        const tfs_contact& c = contacts[i];
        if (c.body_a >= ctx->nb || c.body_b >= ctx->nb) {
            return fail(ctx, TFS_E_ARG, "contact %u names body %u/%u of %u", i, c.body_a, c.body_b, ctx->nb);
        }
        if (c.count < 0 || c.count > 4) { return fail(ctx, TFS_E_ARG, "contact %u has %d points", i, c.count); }
        if (ctx->nWorlds > 1 && ctx->hostBodyWorld[c.body_a] != ctx->hostBodyWorld[c.body_b]) {
            return fail(ctx, TFS_E_ARG, "contact %u joins bodies of two worlds", i);
        }
    }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n > ctx->capC) {
        int r = allocContacts(ctx, n);
        if (r) { return r; }
    }
    if (n) {
        HIP_TRY(hipMemcpy(ctx->contacts, contacts, sizeof(tfs_contact) * n, hipMemcpyHostToDevice));
        HIP_TRY(hipMemset(ctx->ext, 0, sizeof(tfs_contact_ext) * n));
    }
    ctx->nc = n;
    ctx->haveContacts = true;
    return 0;
}

int tfs_upload_contact_ext(tfs_ctx* ctx, const tfs_contact_ext* ext, uint32_t n) {
    if (!ctx || (!ext && n)) { return fail(ctx, TFS_E_ARG, "null argument"); }
    if (n != ctx->nc) { return fail(ctx, TFS_E_ARG, "%u reuse records for %u contacts", n, ctx->nc); }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n) { HIP_TRY(hipMemcpy(ctx->ext, ext, sizeof(tfs_contact_ext) * n, hipMemcpyHostToDevice)); }
    return 0;
}

int tfs_download_contact_ext(tfs_ctx* ctx, tfs_contact_ext* out, uint32_t n) {
    if (!ctx || (!out && n)) { return fail(ctx, TFS_E_ARG, "null argument"); }
    if (n > ctx->nc) { return fail(ctx, TFS_E_ARG, "%u reuse records asked for, %u contacts resident", n, ctx->nc); }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n) { HIP_TRY(hipMemcpy(out, ctx->ext, sizeof(tfs_contact_ext) * n, hipMemcpyDeviceToHost)); }
    return 0;
}

int tfs_stream_contacts(tfs_ctx* ctx, const tfs_pair* pairs, uint32_t n_pairs, const tfs_manifold* manifolds,
                        uint32_t n_manifolds, uint32_t* n_contacts) {
    if (!ctx || (!pairs && n_pairs) || (!manifolds && n_manifolds)) { return fail(ctx, TFS_E_ARG, "null argument"); }
    if (!ctx->haveBodies) { return fail(ctx, TFS_E_STATE, "upload bodies before streaming contacts"); }
    ctx->fixedOn = false;
    const uint32_t nb = ctx->nb, nOld = ctx->haveContacts ? ctx->nc : 0u;
    for (uint32_t p = 0; p < n_pairs; ++p) {
        const tfs_pair& x = pairs[p];
        if (p > 0 && !(pairs[p - 1].key < x.key)) { return fail(ctx, TFS_E_ARG, "pair %u is out of key order", p); }
        if (x.body_a >= nb || x.body_b >= nb) { return fail(ctx, TFS_E_ARG, "pair %u names body %u/%u of %u", p, x.body_a, x.body_b, nb); }
        if ((uint64_t)x.first + x.count > n_manifolds) { return fail(ctx, TFS_E_ARG, "pair %u's manifolds are out of range", p); }
        if (ctx->nWorlds > 1 && ctx->hostBodyWorld[x.body_a] != ctx->hostBodyWorld[x.body_b]) {
            return fail(ctx, TFS_E_ARG, "pair %u joins bodies of two worlds", p);
        }
        for (uint32_t j = 0; j < x.count; ++j) {
            const int32_t c = manifolds[x.first + j].count;
            if (c < 1 || c > 4) { return fail(ctx, TFS_E_ARG, "pair %u's manifold %u has %d points", p, j, c); }
        }
    }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    int r;
    const uint32_t most = nOld + n_manifolds;
    if (most > ctx->streamCap) {
        if ((r = grow(ctx, ctx->contactsNext, most)) || (r = grow(ctx, ctx->extNext, most))) { return r; }
        ctx->streamCap = most;
    }
    if (n_pairs > ctx->pairCap) {
        if ((r = grow(ctx, ctx->pairs, n_pairs)) || (r = grow(ctx, ctx->pairLo, n_pairs)) || (r = grow(ctx, ctx->pairHi, n_pairs)) ||
            (r = grow(ctx, ctx->pairKeep, n_pairs)) || (r = grow(ctx, ctx->pairStart, n_pairs))) {
            return r;
        }
        ctx->pairCap = n_pairs;
    }
    if (n_manifolds > ctx->manifoldCap) {
        if ((r = grow(ctx, ctx->manifolds, n_manifolds))) { return r; }
        ctx->manifoldCap = n_manifolds;
    }
    // the scans' status words cover the larger of the two lists
    if ((size_t)3 * std::max<size_t>(1, ((size_t)std::max(n_pairs, nOld) + kScanTile - 1) / kScanTile) > ctx->statusCap) {
        const uint32_t keepB = ctx->capB;
        ctx->capB = std::max(ctx->capB, std::max(n_pairs, nOld));
        r = ensureScratch(ctx);
        ctx->capB = keepB;
        if (r) { return r; }
    }
    if (n_pairs) { HIP_TRY(hipMemcpyAsync(ctx->pairs, pairs, sizeof(tfs_pair) * n_pairs, hipMemcpyHostToDevice, ctx->stream)); }
    if (n_manifolds) {
        HIP_TRY(hipMemcpyAsync(ctx->manifolds, manifolds, sizeof(tfs_manifold) * n_manifolds, hipMemcpyHostToDevice, ctx->stream));
    }
    HIP_TRY(hipMemsetAsync(ctx->cnt, 0, sizeof(DevCounts), ctx->stream));
    ctx->scanSlot = 0;
    const tfs_settings& s = ctx->settings;
    const float reach = s.contact_reuse_distance * s.length_unit;
    const float reuseCos = std::cos(0.5f * s.contact_reuse_angle);
    const float matchSq = s.speculative * s.speculative;
    if (n_pairs) {
        LAUNCH(k_stream_match, blocksFor(n_pairs), kThreads, ctx->pairs, n_pairs, ctx->contacts, ctx->ext, nOld, ctx->bodies, reach,
               reuseCos, ctx->pairLo, ctx->pairHi, ctx->pairKeep, ctx->pairStart);
    }
    if (nOld) { LAUNCH(k_stream_resting, blocksFor(nOld), kThreads, ctx->contacts, nOld, ctx->bodies, ctx->fActive); }
    const ScanJob jobs[2] = { { ctx->pairStart, nullptr, n_pairs, ctx->pairStart, nullptr, &ctx->cnt->streamPairs },
                              { ctx->fActive, nullptr, nOld, nullptr, ctx->active, &ctx->cnt->streamRest } };
    if ((r = scanJobs(ctx, jobs, 2, std::max(n_pairs, nOld)))) { return r; }
    if (n_pairs) {
        LAUNCH(k_stream_fill, blocksFor(n_pairs), kThreads, ctx->pairs, n_pairs, ctx->pairLo, ctx->pairHi, ctx->pairKeep,
               ctx->pairStart, ctx->manifolds, ctx->contacts, ctx->ext, ctx->active, &ctx->cnt->streamRest, ctx->bodies,
               s.speculative, matchSq, ctx->contactsNext, ctx->extNext);
    }
    if (nOld) {
        LAUNCH(k_stream_rest_fill, blocksFor(nOld), kThreads, ctx->active, &ctx->cnt->streamRest, ctx->contacts, ctx->ext,
               ctx->pairs, n_pairs, ctx->pairStart, &ctx->cnt->streamPairs, ctx->contactsNext, ctx->extNext);
    }
    static_assert(offsetof(DevCounts, streamRest) == offsetof(DevCounts, streamPairs) + sizeof(uint32_t),
                  "the stream's totals are read back together");
    HIP_TRY(hipMemcpyAsync(&ctx->hostCnt->streamPairs, &ctx->cnt->streamPairs, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipMemcpyAsync(&ctx->hostCnt->error, &ctx->cnt->error, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (ctx->hostCnt->error) { return kernelFail(ctx, ctx->hostCnt->error); }
    const uint32_t total = (n_pairs ? ctx->hostCnt->streamPairs : 0u) + (nOld ? ctx->hostCnt->streamRest : 0u);
    if (total > ctx->capC && (r = allocContacts(ctx, total))) { return r; }
    if (total) {
        HIP_TRY(hipMemcpyAsync(ctx->contacts, ctx->contactsNext, sizeof(tfs_contact) * total, hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(ctx->ext, ctx->extNext, sizeof(tfs_contact_ext) * total, hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->nc = total;
    ctx->haveContacts = true;
    if (n_contacts) { *n_contacts = total; }
    return 0;
}

int tfs_step(tfs_ctx* ctx, float dt, tfs_refresh_fn refresh, void* user) {
    if (!ctx) { return TFS_E_ARG; }
    if (!ctx->haveBodies || !ctx->haveContacts) { return fail(ctx, TFS_E_STATE, "upload bodies and contacts first"); }
    if (!(dt > 0.0f) || !(dt <= 3.4e38f)) { return 0; }
    if (hipSetDevice(ctx->device) != hipSuccess) { return fail(ctx, TFS_E_DEVICE, "hipSetDevice failed"); }
    ctx->fixedOn = false;
    return step(ctx, dt, refresh, user);
}

int tfs_set_step_sleep(tfs_ctx* ctx, int on) {
    if (!ctx) { return TFS_E_ARG; }
    ctx->stepSleep = on != 0;
    return 0;
}

int tfs_sleep(tfs_ctx* ctx, float dt) {
    if (!ctx) { return TFS_E_ARG; }
    if (!ctx->haveBodies || !ctx->haveContacts) { return fail(ctx, TFS_E_STATE, "upload bodies and contacts first"); }
    if (!(dt > 0.0f) || !(dt <= 3.4e38f)) { return 0; }
    if (ctx->nJoints && ctx->jointBodies > ctx->nb) {
        return fail(ctx, TFS_E_STATE, "the resident joints name body %u of %u: upload them again", ctx->jointBodies - 1,
                    ctx->nb);
    }
    if (hipSetDevice(ctx->device) != hipSuccess) { return fail(ctx, TFS_E_DEVICE, "hipSetDevice failed"); }
    HIP_TRY(hipMemsetAsync(ctx->cnt, 0, sizeof(DevCounts), ctx->stream));
    int it = 0, r;
    if ((r = launchSleep(ctx, dt, it))) { return r; }
    HIP_TRY(hipMemcpyAsync(&ctx->hostCnt->error, &ctx->cnt->error, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if ((r = finishSleep(ctx, it, false))) { return r; }
    if (ctx->hostCnt->error) { return kernelFail(ctx, ctx->hostCnt->error); }
    return 0;
}

int tfs_step_noalloc(int device, tfs_body* bodies, uint32_t n_bodies, tfs_contact* contacts, uint32_t n_contacts,
                     const tfs_shape* shapes, const tfs_settings* settings, float dt, void* stream) {
    if (device < 0 || device >= kNoAllocDevices || !settings || (n_bodies && !bodies) || (n_contacts && !contacts)) {
        return TFS_E_ARG;
    }
    if (n_bodies > kNoAllocBodies || n_contacts > kNoAllocContacts) { return TFS_E_CAPACITY; }
    if (settings->substeps < 1 || settings->substeps > 64) { return TFS_E_ARG; }
    if (!(dt > 0.0f) || !(dt <= 3.4e38f)) { return 0; }
    if (hipSetDevice(device) != hipSuccess) { return TFS_E_DEVICE; }
    int r;
    if (!g_noallocReady[device] && (r = noallocSetUp(device))) { return r; }
    tfs_ctx* ctx = &g_noalloc[device];
    ctx->stream = (hipStream_t)stream;
    ctx->settings = *settings;
    ctx->nb = n_bodies;
    ctx->nc = n_contacts;
    ctx->nWorlds = 1;
    ctx->haveBodies = ctx->haveContacts = true;
    ctx->fixedOn = false;
    ctx->shapes = const_cast<tfs_shape*>(shapes);
    ctx->nShapes = shapes ? n_bodies : 0;
    // the caller's world into the scratch's arrays, checked before any kernel indexes with it
    if (n_bodies) { HIP_TRY(hipMemcpyAsync(ctx->bodies, bodies, sizeof(tfs_body) * n_bodies, hipMemcpyDeviceToDevice, ctx->stream)); }
    if (n_contacts) {
        HIP_TRY(hipMemcpyAsync(ctx->contacts, contacts, sizeof(tfs_contact) * n_contacts, hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIP_TRY(hipMemsetAsync(ctx->cnt, 0, sizeof(DevCounts), ctx->stream));
    if (n_contacts || ctx->nShapes) {
        LAUNCH(k_check_input, blocksFor(std::max(n_contacts, ctx->nShapes)), kThreads, ctx->contacts, n_contacts, n_bodies,
               ctx->shapes, ctx->nShapes, ctx->cnt);
    }
    if ((r = checkErrors(ctx))) { return r; }
    if ((r = step(ctx, dt, nullptr, nullptr))) { return r; }
    if (n_bodies) { HIP_TRY(hipMemcpyAsync(bodies, ctx->bodies, sizeof(tfs_body) * n_bodies, hipMemcpyDeviceToDevice, ctx->stream)); }
    if (n_contacts) {
        HIP_TRY(hipMemcpyAsync(contacts, ctx->contacts, sizeof(tfs_contact) * n_contacts, hipMemcpyDeviceToDevice, ctx->stream));
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tfs_autotune(tfs_ctx* ctx, uint32_t rounds) {
    if (!ctx) { return TFS_E_ARG; }
    ctx->tuning = false;
    if (rounds == 0) { return 0; }
    ctx->tuneShapes.clear();
    for (uint32_t blocks : { 64u, 128u, 256u }) {
        if (blocks > ctx->passBlockLimit) { continue; }
        for (uint32_t tail : { 0u, 64u, 128u }) {
            for (bool made : { false, true }) { ctx->tuneShapes.push_back(tfs_ctx::PassShape{ blocks, tail, made, false }); }
        }
        if (blocks <= ctx->fusedBlockLimit) {
            for (bool made : { false, true }) { ctx->tuneShapes.push_back(tfs_ctx::PassShape{ blocks, 64u, made, true }); }
        }
    }
    if (ctx->tuneShapes.empty()) { ctx->tuneShapes.push_back(tfs_ctx::PassShape{ ctx->passBlockLimit, 64u, false, false }); }
    ctx->tuneMs.assign(ctx->tuneShapes.size(), std::vector<float>());
    ctx->tuneStep = 0;
    ctx->tuneRounds = rounds;
    ctx->tuning = true;
    return 0;
}

int tfs_pass_shape(tfs_ctx* ctx, uint32_t shape[7]) {
    if (!ctx || !shape) { return TFS_E_ARG; }
    shape[0] = ctx->maxPassBlocks;
    shape[1] = ctx->passTail ? ctx->tailSegRows : 0u;
    shape[2] = ctx->madeAxes ? 1u : 0u;
    shape[3] = ctx->tuning ? (uint32_t)(ctx->tuneRounds * ctx->tuneShapes.size() - ctx->tuneStep) : 0u;
    shape[4] = ctx->tuning ? (uint32_t)ctx->tuneShapes.size() : 0u;
    shape[5] = ctx->fused ? 1u : 0u;
    shape[6] = ctx->useGraph ? 1u : 0u;
    return 0;
}

int tfs_set_pass_shape(tfs_ctx* ctx, uint32_t blocks, uint32_t tail_rows, int made_axes, int fused) {
    if (!ctx) { return TFS_E_ARG; }
    if (blocks == 0 || blocks > ctx->passBlockLimit) {
        return fail(ctx, TFS_E_ARG, "blocks must be in 1..%u (the blocks resident at once)", ctx->passBlockLimit);
    }
    if (tail_rows > (uint32_t)kPassThreads) { return fail(ctx, TFS_E_ARG, "tail_rows must be at most %d", kPassThreads); }
    if (fused && blocks > ctx->fusedBlockLimit) {
        return fail(ctx, TFS_E_ARG, "a fused substep takes at most %u blocks", ctx->fusedBlockLimit);
    }
    ctx->tuning = false;
    applyShape(ctx, tfs_ctx::PassShape{ blocks, tail_rows, made_axes != 0, fused != 0 });
    return 0;
}

int tfs_set_graph(tfs_ctx* ctx, int on) {
    if (!ctx) { return TFS_E_ARG; }
    ctx->useGraph = on != 0;
    return 0;
}

int tfs_set_cooperative(tfs_ctx* ctx, int on) {
    if (!ctx) { return TFS_E_ARG; }
    ctx->coopLaunch = on != 0;
    return 0;
}

int tfs_set_body_shapes(tfs_ctx* ctx, uint32_t n, const tfs_shape* shapes) {
    if (!ctx || (n && !shapes)) { return TFS_E_ARG; }
    for (uint32_t i = 0; i < n; ++i) {
        const tfs_shape& x = shapes[i];
        const bool ok = x.kind == TFS_SHAPE_HALF_SPACE ||
                        (x.kind == TFS_SHAPE_SPHERE && x.radius > 0.0f && x.radius <= 3.4e38f) ||
                        (x.kind == TFS_SHAPE_BOX && x.half[0] > 0.0f && x.half[0] <= 3.4e38f && x.half[1] > 0.0f &&
                         x.half[1] <= 3.4e38f && x.half[2] > 0.0f && x.half[2] <= 3.4e38f);
        if (!ok) { return fail(ctx, TFS_E_ARG, "shape %u: kind %u with these sizes is not a sphere, box or half-space", i, x.kind); }
    }
    if (hipSetDevice(ctx->device) != hipSuccess) { return fail(ctx, TFS_E_DEVICE, "hipSetDevice failed"); }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n > ctx->shapesCap) {
        int r;
        if ((r = grow(ctx, ctx->shapes, n))) { return r; }
        ctx->shapesCap = n;
    }
    if (n) { HIP_TRY(hipMemcpy(ctx->shapes, shapes, sizeof(tfs_shape) * n, hipMemcpyHostToDevice)); }
    ctx->nShapes = n;
    return 0;
}

int tfs_upload_joints(tfs_ctx* ctx, const tfs_joint* joints, uint32_t n, const float* world_motion) {
    if (!ctx || (n && !joints)) { return TFS_E_ARG; }
    if (ctx->arena) { return fail(ctx, TFS_E_STATE, "a context that may not allocate takes no joints"); }
    uint32_t named = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const tfs_joint& j = joints[k];
        named = std::max(named, std::max(j.body_b, j.body_a == TFS_WORLD_BODY ? 0u : j.body_a) + 1);
        if (j.body_b >= ctx->nb || (j.body_a != TFS_WORLD_BODY && (j.body_a >= ctx->nb || j.body_a == j.body_b))) {
            return fail(ctx, TFS_E_ARG, "joint %u: bodies %u and %u out of range for %u bodies", k, j.body_a, j.body_b, ctx->nb);
        }
        if (j.alive && !tfj::supported(j)) {
            return fail(ctx, TFS_E_ARG, "joint %u: a %s reads angles through atan2; the device does not solve it", k,
                        j.type == TFS_JOINT_HINGE ? "limited hinge" : "swing-twist joint");
        }
    }
    if (hipSetDevice(ctx->device) != hipSuccess) { return fail(ctx, TFS_E_DEVICE, "hipSetDevice failed"); }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    int r;
    if (!ctx->worldMotion && (r = grow(ctx, ctx->worldMotion, 2))) { return r; }
    if (n > ctx->capJ) {
        if ((r = grow(ctx, ctx->joints, n)) || (r = grow(ctx, ctx->jointActive, n)) ||
            (r = grow(ctx, ctx->edges, (size_t)ctx->capC + n))) {
            return r;
        }
        ctx->capJ = n;
    }
    if (n) { HIP_TRY(hipMemcpy(ctx->joints, joints, sizeof(tfs_joint) * n, hipMemcpyHostToDevice)); }
    float4 wm[2] = { make_float4(0.0f, 0.0f, 0.0f, 0.0f), make_float4(0.0f, 0.0f, 0.0f, 0.0f) };
    if (world_motion) {
        wm[0] = make_float4(world_motion[0], world_motion[1], world_motion[2], 0.0f);
        wm[1] = make_float4(world_motion[3], world_motion[4], world_motion[5], 0.0f);
    }
    HIP_TRY(hipMemcpy(ctx->worldMotion, wm, sizeof wm, hipMemcpyHostToDevice));
    ctx->nJoints = n;
    ctx->jointBodies = named;
    return 0;
}

int tfs_download_joints(tfs_ctx* ctx, tfs_joint* out, uint32_t n, float* world_motion) {
    if (!ctx || (n && !out)) { return TFS_E_ARG; }
    if (n > ctx->nJoints) { return fail(ctx, TFS_E_ARG, "%u joints asked for, %u resident", n, ctx->nJoints); }
    if (hipSetDevice(ctx->device) != hipSuccess) { return fail(ctx, TFS_E_DEVICE, "hipSetDevice failed"); }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n) { HIP_TRY(hipMemcpy(out, ctx->joints, sizeof(tfs_joint) * n, hipMemcpyDeviceToHost)); }
    if (world_motion) {
        float4 wm[2] = { make_float4(0.0f, 0.0f, 0.0f, 0.0f), make_float4(0.0f, 0.0f, 0.0f, 0.0f) };
        if (ctx->worldMotion) { HIP_TRY(hipMemcpy(wm, ctx->worldMotion, sizeof wm, hipMemcpyDeviceToHost)); }
        const float v[6] = { wm[0].x, wm[0].y, wm[0].z, wm[1].x, wm[1].y, wm[1].z };
        std::memcpy(world_motion, v, sizeof v);
    }
    return 0;
}

int tfs_fixed_begin(tfs_ctx* ctx, float dt, uint32_t max_steps) {
    if (!ctx) { return TFS_E_ARG; }
    if (!ctx->haveBodies || !ctx->haveContacts) { return fail(ctx, TFS_E_STATE, "upload bodies and contacts first"); }
    if (!(dt > 0.0f) || !(dt <= 3.4e38f) || max_steps == 0) { return fail(ctx, TFS_E_ARG, "dt and max_steps must be positive"); }
    if (hipSetDevice(ctx->device) != hipSuccess) { return fail(ctx, TFS_E_DEVICE, "hipSetDevice failed"); }
    return fixedBegin(ctx, dt, max_steps);
}

int tfs_fixed_step(tfs_ctx* ctx) {
    if (!ctx) { return TFS_E_ARG; }
    if (!ctx->fixedOn) { return fail(ctx, TFS_E_STATE, "call tfs_fixed_begin first"); }
    if (ctx->fixedSteps >= ctx->fixedMax) { return fail(ctx, TFS_E_STATE, "all %u checkpoints are taken", ctx->fixedMax); }
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t nr = ctx->fixedCounts.colorStart[TFS_COLORS];
    LAUNCH(k_fx_save, blocksFor(std::max(ctx->nb, nr)), kThreads, ctx->bodies, ctx->nb, ctx->motion, ctx->rows, nr,
           checkpointOf(ctx, ctx->fixedSteps));
    int r;
    if ((r = fixedStep(ctx, false))) { return r; }
    ++ctx->fixedSteps;
    ctx->fixedLeft = 0;
    if ((r = checkErrors(ctx))) { return r; }
    return collectProfile(ctx);
}

int tfs_adjoint_seed(tfs_ctx* ctx, const float* dl_dcom) {
    if (!ctx || !dl_dcom) { return TFS_E_ARG; }
    if (!ctx->fixedOn || ctx->fixedSteps == 0) { return fail(ctx, TFS_E_STATE, "take fixed-contact steps first"); }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipMemcpyAsync(ctx->dlcom, dl_dcom, 3 * sizeof(float) * ctx->nb, hipMemcpyHostToDevice, ctx->stream));
    const uint32_t nr = ctx->fixedCounts.colorStart[TFS_COLORS];
    LAUNCH(k_adj_seed, blocksFor(std::max(ctx->nb, nr)), kThreads, ctx->nb, ctx->dlcom, nr, ctx->rows.cap, ctx->adj);
    ctx->fixedLeft = ctx->fixedSteps;
    return checkErrors(ctx);
}

int tfs_adjoint_step(tfs_ctx* ctx) {
    if (!ctx) { return TFS_E_ARG; }
    if (!ctx->fixedOn || ctx->fixedLeft == 0) { return fail(ctx, TFS_E_STATE, "no recorded step left to reverse"); }
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t nb = ctx->nb, nr = ctx->fixedCounts.colorStart[TFS_COLORS];
    const uint32_t k = ctx->fixedLeft - 1;
    const Ctx c = makeCtx(ctx->settings, ctx->fixedDt);
    const int n = ctx->settings.substeps > 1 ? ctx->settings.substeps : 1;
    // the step again from its checkpoint, recording what each row saw
    LAUNCH(k_fx_restore, blocksFor(std::max(nb, nr)), kThreads, ctx->bodies, nb, ctx->motion, ctx->rows, nr,
           checkpointOf(ctx, k));
    int r;
    if ((r = fixedStep(ctx, true))) { return r; }
    // then backwards through it
    LAUNCH(k_adj_rebase, blocksFor(nb), kThreads, nb, ctx->bdeg, ctx->bstart, ctx->bep, ctx->rows, ctx->adj);
    for (int i = n - 1; i >= 0; --i) {
        if ((r = passAdjoint(ctx, kRelax, 2 * i + 1, c))) { return r; }
        LAUNCH(k_adj_positions, blocksFor(nb), kThreads, ctx->bodies, nb, c.h, ctx->adj);
        if ((r = passAdjoint(ctx, kBias, 2 * i, c))) { return r; }
        LAUNCH(k_adj_warm, blocksFor(nr), kThreads, ctx->rows, nr, ctx->adj);
    }
    ctx->fixedLeft = k;
    if ((r = checkErrors(ctx))) { return r; }
    return collectProfile(ctx);
}

int tfs_adjoint_read(tfs_ctx* ctx, float* dl_dv, float* dl_dw, float* dl_dfriction) {
    if (!ctx) { return TFS_E_ARG; }
    if (!ctx->fixedOn) { return fail(ctx, TFS_E_STATE, "call tfs_fixed_begin first"); }
    HIP_TRY(hipSetDevice(ctx->device));
    const uint32_t nb = ctx->nb, nr = ctx->fixedCounts.colorStart[TFS_COLORS];
    if (dl_dfriction && ctx->nc) {
        HIP_TRY(hipMemsetAsync(ctx->muContact, 0, sizeof(float) * ctx->nc, ctx->stream));
        LAUNCH(k_adj_friction, blocksFor(nr), kThreads, ctx->rows, nr, ctx->adj.muBar, ctx->muContact);
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (dl_dv || dl_dw) {
        std::vector<float4> bar(2 * (size_t)nb);
        if (nb) { HIP_TRY(hipMemcpy(bar.data(), ctx->adj.bar, sizeof(float4) * 2 * nb, hipMemcpyDeviceToHost)); }
        for (uint32_t i = 0; i < nb; ++i) {
            const float4 v = bar[2 * i], w = bar[2 * i + 1];
            if (dl_dv) {
                dl_dv[3 * i] = v.x;
                dl_dv[3 * i + 1] = v.y;
                dl_dv[3 * i + 2] = v.z;
            }
            if (dl_dw) {
                dl_dw[3 * i] = w.x;
                dl_dw[3 * i + 1] = w.y;
                dl_dw[3 * i + 2] = w.z;
            }
        }
    }
    if (dl_dfriction && ctx->nc) {
        HIP_TRY(hipMemcpy(dl_dfriction, ctx->muContact, sizeof(float) * ctx->nc, hipMemcpyDeviceToHost));
    }
    return 0;
}

int tfs_download_poses(tfs_ctx* ctx, tfs_pose* out, uint32_t n) {
    if (!ctx || (!out && n)) { return fail(ctx, TFS_E_ARG, "null argument"); }
    if (n > ctx->nb) { return fail(ctx, TFS_E_ARG, "%u poses asked for, %u bodies resident", n, ctx->nb); }
    HIP_TRY(hipSetDevice(ctx->device));
    if (n == 0) { return 0; }
    // packed on the device into the contact-sized staging array, then one copy
    const size_t need = sizeof(tfs_pose) * n;
    if (need > sizeof(tfs_contact) * (size_t)ctx->capC) {
        int r;
        if ((r = allocContacts(ctx, (uint32_t)((need + sizeof(tfs_contact) - 1) / sizeof(tfs_contact))))) { return r; }
    }
    tfs_pose* staged = (tfs_pose*)ctx->fresh;
    LAUNCH(k_poses, blocksFor(n), kThreads, ctx->bodies, n, staged);
    HIP_TRY(hipMemcpyAsync(out, staged, need, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return 0;
}

int tfs_download_bodies(tfs_ctx* ctx, tfs_body* out, uint32_t n) {
    if (!ctx || (!out && n)) { return fail(ctx, TFS_E_ARG, "null argument"); }
    if (n > ctx->nb) { return fail(ctx, TFS_E_ARG, "asked for %u bodies of %u", n, ctx->nb); }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n) { HIP_TRY(hipMemcpy(out, ctx->bodies, sizeof(tfs_body) * n, hipMemcpyDeviceToHost)); }
    return 0;
}

int tfs_download_contacts(tfs_ctx* ctx, tfs_contact* out, uint32_t n) {
    if (!ctx || (!out && n)) { return fail(ctx, TFS_E_ARG, "null argument"); }
    if (n > ctx->nc) { return fail(ctx, TFS_E_ARG, "asked for %u contacts of %u", n, ctx->nc); }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n) { HIP_TRY(hipMemcpy(out, ctx->contacts, sizeof(tfs_contact) * n, hipMemcpyDeviceToHost)); }
    return 0;
}

int tfs_download_list(tfs_ctx* ctx, int kind, void* out, uint32_t cap, uint32_t* count) {
    if (!ctx || !count) { return fail(ctx, TFS_E_ARG, "null argument"); }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const DevCounts& h = *ctx->hostCnt;
    const void* src = nullptr;
    uint32_t n = 0;
    size_t size = sizeof(uint32_t);
    switch (kind) {
        case 0: src = ctx->active; n = h.nActive; break;
        case 1: src = ctx->fast; n = h.nFast; break;
        case 2: src = ctx->fastSolid; n = h.nFastSolid; break;
        case 3: src = ctx->impact; n = h.nImpact; break;
        case 4: src = ctx->rowsList; n = h.nActive; break;
        case 5: src = ctx->impactRowsList; n = h.nImpact; break;
        case 6: src = ctx->colors; n = ctx->nc; size = 1; break;
        default: return fail(ctx, TFS_E_ARG, "unknown list kind %d", kind);
    }
    *count = n;
    const uint32_t m = n < cap ? n : cap;
    if (m && out) { HIP_TRY(hipMemcpy(out, src, size * m, hipMemcpyDeviceToHost)); }
    return 0;
}

int tfs_download_touches(tfs_ctx* ctx, uint32_t* events, uint32_t events_cap, uint32_t* n_events, tfs_hit* hits,
                         uint32_t hits_cap, uint32_t* n_hits, uint32_t* touching) {
    if (!ctx) { return TFS_E_ARG; }
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const uint32_t ne = ctx->touchCounts[0], nh = ctx->touchCounts[1];
    if (events && ne) { HIP_TRY(hipMemcpy(events, ctx->touchList, sizeof(uint32_t) * std::min(ne, events_cap), hipMemcpyDeviceToHost)); }
    if (hits && nh) { HIP_TRY(hipMemcpy(hits, ctx->hits, sizeof(tfs_hit) * std::min(nh, hits_cap), hipMemcpyDeviceToHost)); }
    if (n_events) { *n_events = ne; }
    if (n_hits) { *n_hits = nh; }
    if (touching) { *touching = ctx->touchCounts[2]; }
    return 0;
}

int tfs_last_step_ms(tfs_ctx* ctx, float* solveMs, float* sleepMs) {
    if (!ctx) { return TFS_E_ARG; }
    if (solveMs) { *solveMs = ctx->solveMs; }
    if (sleepMs) { *sleepMs = ctx->sleepMs; }
    return 0;
}

int tfs_set_profiling(tfs_ctx* ctx, int on) {
    if (!ctx) { return TFS_E_ARG; }
    ctx->profiling = on != 0;
    if (!on) { return 0; }
    ctx->pending.clear();
    for (int f = 0; f < kProfileFamilies; ++f) {
        ctx->profLaunches[f] = 0;
        ctx->profMs[f] = 0.0;
        ctx->profBytes[f] = 0.0;
    }
    return 0;
}

int tfs_read_profile(tfs_ctx* ctx, int family, uint64_t* launches, double* ms, double* bytes) {
    if (!ctx || family < 0 || family >= kProfileFamilies) { return fail(ctx, TFS_E_ARG, "unknown profile family"); }
    if (launches) { *launches = ctx->profLaunches[family]; }
    if (ms) { *ms = ctx->profMs[family]; }
    if (bytes) { *bytes = ctx->profBytes[family]; }
    return 0;
}

int tfs_last_step_counts(tfs_ctx* ctx, uint32_t counts[8]) {
    if (!ctx || !counts) { return TFS_E_ARG; }
    memcpy(counts, ctx->lastCounts, sizeof ctx->lastCounts);
    return 0;
}

}  // extern "C"
