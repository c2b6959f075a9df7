// nova-mi355x: MDP value iteration on one MI355X (gfx950, wave64).
//
// The sweep kernel reproduces src/mdp/algorithms/mdp_vi_cpu.cpp:36-63 bit for bit:
//   Q(s,a) = R[s,a]; Q += (gamma*T[s,a,i]) * V[S[s,a,i]] for i = 0.. until the first S < 0;
//   V'(s) = Q(s,0) unless a later action is strictly larger; pi(s) = that action.
// Layout and mapping (one workgroup = 256 lanes = 4 waves per tile of whole states):
//   * each (s,a) successor row of ns entries is owned by G adjacent lanes holding E = 4 (or 1)
//     consecutive entries each, so a wave reads 64*E*4 contiguous bytes of S and of T per load
//     (dwordx4, non-temporal: S/T are streamed once per sweep and must not evict V from L2);
//   * the valid prefix of each row (up to the first negative successor) is found with a wave
//     ballot before any gather, so V is only read at valid indices;
//   * the row sum runs as a serial chain across the G lanes (lane g adds its E products in
//     order, then hands the running sum to lane g+1), keeping the reference's left-to-right
//     evaluation order and its separate rounding of each product and sum (no FMA);
//   * row sums are staged in LDS; one lane per state then takes the max over actions with the
//     reference's tie rule, and V'/pi are written coalesced.
// A plain one-thread-per-state kernel covers shapes the tiled kernel does not (ns > 256,
// m > 2048, misaligned arrays).
#include <cstdlib>
#include <cstring>
#include <vector>

#include <hipcub/hipcub.hpp>

#include <nova/amd/nova_amd.h>
#include <nova/mdp/algorithms/mdp_vi_gpu.h>

#include "nova_hip.h"

namespace nova_amd {

namespace {

constexpr int kBlock = 256;
constexpr int kTileRows = 2048;   // (s,a) rows whose Q values a workgroup stages in LDS
constexpr int kUnroll = 8;        // rows in flight per lane group (8 measured ~4% faster than 4)

typedef int int4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

template <int E> struct Entries {
    int s[E];
    float t[E];
};

template <int E>
__device__ __forceinline__ void load_entries(const int *__restrict__ S, const float *__restrict__ T,
                                             size_t off, bool in, Entries<E> &x)
{
    if constexpr (E == 4) {
        if (in) {
            int4v sv = __builtin_nontemporal_load(reinterpret_cast<const int4v *>(S + off));
            float4v tv = __builtin_nontemporal_load(reinterpret_cast<const float4v *>(T + off));
            x.s[0] = sv.x; x.s[1] = sv.y; x.s[2] = sv.z; x.s[3] = sv.w;
            x.t[0] = tv.x; x.t[1] = tv.y; x.t[2] = tv.z; x.t[3] = tv.w;
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++) { x.s[e] = -1; x.t[e] = 0.0f; }
        }
    } else {
        x.s[0] = in ? __builtin_nontemporal_load(S + off) : -1;
        x.t[0] = in ? __builtin_nontemporal_load(T + off) : 0.0f;
    }
}

template <int E, int G>
__global__ __launch_bounds__(kBlock) void vi_sweep_tiled(
        unsigned n, unsigned m, unsigned ns, unsigned statesPerTile, float gamma,
        const int *__restrict__ S, const float *__restrict__ T, const float *__restrict__ R,
        const float *__restrict__ V, float *__restrict__ Vp, unsigned *__restrict__ pi,
        int writePi)
{
    __shared__ float q[kTileRows];

    const unsigned s0 = blockIdx.x * statesPerTile;
    const unsigned s1 = min(n, s0 + statesPerTile);
    const unsigned rows = (s1 - s0) * m;
    const size_t rowBase = (size_t)s0 * m;

    const unsigned g = threadIdx.x % G;                 // lane within its row group
    const unsigned lane = threadIdx.x % 64;             // lane within the wave
    const unsigned long long groupBelow =               // lanes of my group before me
        ((1ull << g) - 1ull) << (lane - g);
    const unsigned passRows = kBlock / G;
    const bool inRow = g * E < ns;                      // my entries lie inside the row
    const unsigned nClamp = n - 1;

    for (unsigned base = threadIdx.x / G; base < rows; base += passRows * kUnroll) {
        Entries<E> x[kUnroll];
        float r0[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const unsigned rr = base + u * passRows;
            const bool live = rr < rows;
            const size_t row = rowBase + rr;
            load_entries<E>(S, T, row * ns + g * E, live && inRow, x[u]);
            r0[u] = (live && g == 0) ? __builtin_nontemporal_load(R + row) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const unsigned rr = base + u * passRows;
            // valid entries of my slice: up to my first negative successor, none if an
            // earlier lane of the row already hit one
            int loc = E;
#pragma unroll
            for (int e = E - 1; e >= 0; e--)
                if (x[u].s[e] < 0) loc = e;
            const unsigned long long hit = __ballot(loc < E);
            const int valid = (hit & groupBelow) ? 0 : loc;
            float p[E];
#pragma unroll
            for (int e = 0; e < E; e++) {
                p[e] = 0.0f;
                if (e < valid) {
                    const unsigned sp = min((unsigned)x[u].s[e], nClamp);
                    p[e] = __fmul_rn(__fmul_rn(gamma, x[u].t[e]), V[sp]);
                }
            }
            // serial chain: lane 0 of the row starts from R, each lane adds its products in
            // order, then passes the running sum one lane up
            float acc = r0[u];
#pragma unroll
            for (int t = 0; t < G; t++) {
                if ((int)g == t) {
#pragma unroll
                    for (int e = 0; e < E; e++)
                        if (e < valid) acc = __fadd_rn(acc, p[e]);
                }
                if (t + 1 < G) {
                    const float up = __shfl_up(acc, 1, G);
                    if ((int)g == t + 1) acc = up;
                }
            }
            if (g == G - 1 && rr < rows)
                q[rr] = acc;
        }
    }
    __syncthreads();

    for (unsigned t = threadIdx.x; t < s1 - s0; t += kBlock) {
        const float *qs = q + (size_t)t * m;
        float best = qs[0];
        unsigned bestA = 0;
        for (unsigned a = 1; a < m; a++) {
            if (qs[a] > best) {
                best = qs[a];
                bestA = a;
            }
        }
        Vp[s0 + t] = best;
        if (writePi)
            pi[s0 + t] = bestA;
    }
}

// One thread per state, sequential: the shapes the tiled kernel does not take.
__global__ __launch_bounds__(kBlock) void vi_sweep_plain(
        unsigned n, unsigned m, unsigned ns, float gamma,
        const int *__restrict__ S, const float *__restrict__ T, const float *__restrict__ R,
        const float *__restrict__ V, float *__restrict__ Vp, unsigned *__restrict__ pi,
        int writePi)
{
    const unsigned s = blockIdx.x * kBlock + threadIdx.x;
    if (s >= n)
        return;
    float best = 0.0f;
    unsigned bestA = 0;
    for (unsigned a = 0; a < m; a++) {
        const size_t row = ((size_t)s * m + a) * ns;
        float q = R[(size_t)s * m + a];
        for (unsigned i = 0; i < ns; i++) {
            const int sp = S[row + i];
            if (sp < 0)
                break;
            q = __fadd_rn(q, __fmul_rn(__fmul_rn(gamma, T[row + i]), V[min((unsigned)sp, n - 1)]));
        }
        if (a == 0 || q > best) {
            best = q;
            bestA = a;
        }
    }
    Vp[s] = best;
    if (writePi)
        pi[s] = bestA;
}

typedef void (*TiledFn)(unsigned, unsigned, unsigned, unsigned, float, const int *,
                        const float *, const float *, const float *, float *, unsigned *, int);

template <int E> TiledFn tiled_for(unsigned G)
{
    switch (G) {
    case 1: return vi_sweep_tiled<E, 1>;
    case 2: return vi_sweep_tiled<E, 2>;
    case 4: return vi_sweep_tiled<E, 4>;
    case 8: return vi_sweep_tiled<E, 8>;
    case 16: return vi_sweep_tiled<E, 16>;
    case 32: return vi_sweep_tiled<E, 32>;
    case 64: return vi_sweep_tiled<E, 64>;
    default: return nullptr;
    }
}

// ---- successor-sorted tiles (large models)
//
// The sweep above is bound by L2 requests: each of its n*m*ns gathers V[S[s,a,i]] is one
// request, since lanes of a wave hit unrelated lines. A one-time layout removes a good part
// of them: the entries of each tile of kSortedEntries consecutive (s, a, i) slots are stored
// sorted by successor, with their T and their slot in the tile. A workgroup streams its tile
// in that order -- neighbouring gathers now share cache lines within a wave instruction and
// across consecutive ones -- and scatters each product (gamma*T)*V into an LDS array at its
// slot; every row is then summed from LDS in its own order, so the sums and V' are the same
// bit for bit as above. Entries past a row's first negative successor are left out.
constexpr unsigned kSortedEntries = 16384;   // slots per tile: two workgroups per CU fit LDS
constexpr int kSortedThreads = 1024;
constexpr int kSortedBatch = 8;             // entries per lane per round of gathers

__global__ __launch_bounds__(256) void vi_row_counts(size_t rows, unsigned ns,
        const int *__restrict__ S, unsigned char *__restrict__ cnt)
{
    const size_t row = blockIdx.x * (size_t)256 + threadIdx.x;
    if (row >= rows)
        return;
    unsigned c = 0;
    while (c < ns && S[row * ns + c] >= 0)
        c++;
    cnt[row] = (unsigned char)(c > 255 ? 255 : c);   // ns <= 255 for this layout
}

// key = successor (clamped below n) for a valid slot, n past a row's end or past the model;
// value = the slot within its tile, numbered entry-major (i * rowsPerTile + local row) so that
// the row sums read their products from LDS with consecutive lanes on consecutive words
__global__ __launch_bounds__(256) void vi_sort_keys(size_t total, unsigned n, unsigned m,
        unsigned ns, unsigned statesPerTile, unsigned tileEntries, const int *__restrict__ S,
        const unsigned char *__restrict__ cnt, int *__restrict__ keys,
        unsigned *__restrict__ vals)
{
    const size_t x = blockIdx.x * (size_t)256 + threadIdx.x;
    if (x >= total)
        return;
    const size_t tile = x / tileEntries;
    const unsigned e = (unsigned)(x % tileEntries);
    const unsigned rowLen = m * ns;
    const size_t s = tile * statesPerTile + e / rowLen;
    const unsigned rem = e % rowLen, i = rem % ns;
    int key = (int)n;
    if (s < n && (e / rowLen) < statesPerTile && i < cnt[s * m + rem / ns])
        key = min(S[s * rowLen + rem], (int)n - 1);
    keys[x] = key;
    vals[x] = i * (statesPerTile * m) + (e / rowLen) * m + rem / ns;
}

// the permuted T and 16-bit slots, and the valid count of every tile
__global__ __launch_bounds__(256) void vi_sort_finish(size_t total, unsigned n, unsigned m,
        unsigned ns, unsigned statesPerTile, unsigned tileEntries, const float *__restrict__ T,
        const int *__restrict__ sortedS, const unsigned *__restrict__ vals,
        float *__restrict__ sortedT, unsigned short *__restrict__ slot,
        unsigned *__restrict__ tileValid)
{
    const size_t x = blockIdx.x * (size_t)256 + threadIdx.x;
    if (x >= total)
        return;
    const size_t tile = x / tileEntries;
    const unsigned e = vals[x];                       // i * rowsPerTile + local row
    const unsigned rowsPerTile = statesPerTile * m;
    const unsigned i = e / rowsPerTile, row = e % rowsPerTile;
    const size_t s = tile * statesPerTile + row / m;
    const bool valid = sortedS[x] < (int)n;
    sortedT[x] = valid ? T[(s * m + row % m) * ns + i] : 0.0f;
    slot[x] = (unsigned short)e;
    // the last valid entry of a tile records the count (keys n sort after every valid one)
    const bool lastValid = valid && ((x + 1) % tileEntries == 0 || sortedS[x + 1] >= (int)n);
    if (lastValid)
        tileValid[tile] = (unsigned)(x % tileEntries) + 1;
}

__global__ __launch_bounds__(kSortedThreads) void vi_sweep_sorted(unsigned n, unsigned m,
        unsigned ns, unsigned statesPerTile, unsigned tileEntries, float gamma,
        const int *__restrict__ sortedS, const float *__restrict__ sortedT,
        const unsigned short *__restrict__ slot, const unsigned *__restrict__ tileValid,
        const unsigned char *__restrict__ cnt, const float *__restrict__ R,
        const float *__restrict__ V, float *__restrict__ Vp, unsigned *__restrict__ pi,
        int writePi)
{
    extern __shared__ float lds[];
    float *P = lds;                       // tileEntries products, by slot (entry-major)
    float *Q = lds + tileEntries;         // statesPerTile * m row sums
    const unsigned tile = blockIdx.x, t = threadIdx.x;
    const size_t base = (size_t)tile * tileEntries;
    const unsigned valid = tileValid[tile];
    // kSortedBatch entries per lane in flight: all stream loads, then all gathers, then the
    // LDS stores (a tile takes two or three such rounds)
    for (unsigned x0 = t; x0 < valid; x0 += kSortedBatch * kSortedThreads) {
        int sp[kSortedBatch];
        float tt[kSortedBatch];
        unsigned short sl[kSortedBatch];
#pragma unroll
        for (int u = 0; u < kSortedBatch; u++) {
            const unsigned x = x0 + u * kSortedThreads;
            const bool in = x < valid;
            sp[u] = in ? __builtin_nontemporal_load(sortedS + base + x) : 0;
            tt[u] = in ? __builtin_nontemporal_load(sortedT + base + x) : 0.0f;
            sl[u] = in ? slot[base + x] : 0;
        }
        float v[kSortedBatch];
#pragma unroll
        for (int u = 0; u < kSortedBatch; u++)
            v[u] = V[sp[u]];
#pragma unroll
        for (int u = 0; u < kSortedBatch; u++)
            if (x0 + u * kSortedThreads < valid)
                P[sl[u]] = __fmul_rn(__fmul_rn(gamma, tt[u]), v[u]);
    }
    __syncthreads();
    const unsigned s0 = tile * statesPerTile;
    const unsigned states = min(statesPerTile, n - s0);
    const unsigned rows = states * m;
    for (unsigned row = t; row < rows; row += kSortedThreads) {
        const size_t g = (size_t)s0 * m + row;
        float q = R[g];
        const unsigned c = cnt[g];
        const unsigned rowsPerTile = statesPerTile * m;
        for (unsigned i = 0; i < c; i++)
            q = __fadd_rn(q, P[i * rowsPerTile + row]);
        Q[row] = q;
    }
    __syncthreads();
    for (unsigned x = t; x < states; x += kSortedThreads) {
        const float *qs = Q + (size_t)x * m;
        float best = qs[0];
        unsigned bestA = 0;
        for (unsigned a = 1; a < m; a++) {
            if (qs[a] > best) {
                best = qs[a];
                bestA = a;
            }
        }
        Vp[s0 + x] = best;
        if (writePi)
            pi[s0 + x] = bestA;
    }
}

// ---- successor-chunked layout: V chunks in LDS, products through HBM (large models)
//
// The sorted-tile sweep above still gathers V from L2: about one L2 request per two entries.
// This layout takes the gathers off L2 altogether. The successor space is cut into chunks of
// kChunkStates states (V of a chunk fills LDS), and every valid (s, a, i) entry is stored in
// the segment of its (successor chunk c, state tile t), segments ordered chunk-major:
//   pass A (vi_chunk_products): a workgroup loads one chunk of V into LDS and streams a
//     contiguous run of that chunk's entries -- the successor's 16-bit offset in the chunk and
//     T -- writing each product (gamma*T)*V[sp] to P at the entry's own position, coalesced;
//   pass B (vi_chunk_rows): a workgroup per state tile reads its segment of every chunk, P and
//     the entry's 16-bit slot in the tile, places the products into LDS by slot, and sums every
//     row in its own order from R, then takes the max over actions.
// Both passes stream; every V read is an LDS read. The products and the order of each row's
// sums are those of the sweeps above, so V' and pi are the same bit for bit.
// Per valid entry: 2 + 4 bytes read and 4 written by pass A, 4 + 2 read by pass B.
constexpr unsigned kChunkStatesDefault = 32768;    // 128 KB of V per pass-A workgroup
constexpr unsigned kPieceEntriesDefault = 32768;   // entries per pass-A workgroup
constexpr unsigned kChunkTileDefault = 8192;       // slots per pass-B tile
constexpr unsigned kChunkAlign = 64;               // entries: chunk starts and piece lengths
constexpr int kChunkThreads = 1024;
constexpr int kChunkVec = 8;                       // entries per lane per load round (pass A)

// key = group * tiles + tile of a valid entry, where group = part * chunks + chunk (part: the
// tile's run of tilesPerPart tiles), and groups * tiles for a slot past a row's valid prefix or
// past n; value = the slot's flat index over the tiles (state-major within a tile)
__global__ __launch_bounds__(256) void vi_chunk_keys(size_t total, unsigned n, unsigned m,
        unsigned ns, unsigned statesPerTile, unsigned tileEntries, unsigned chunkStates,
        unsigned tiles, unsigned chunks, unsigned tilesPerPart, unsigned groups,
        const int *__restrict__ S, const unsigned char *__restrict__ cnt,
        unsigned *__restrict__ keys, unsigned *__restrict__ vals)
{
    const size_t x = blockIdx.x * (size_t)256 + threadIdx.x;
    if (x >= total)
        return;
    const unsigned tile = (unsigned)(x / tileEntries);
    const unsigned e = (unsigned)(x % tileEntries);
    const unsigned rowLen = m * ns;
    const size_t s = (size_t)tile * statesPerTile + e / rowLen;
    const unsigned rem = e % rowLen, i = rem % ns;
    unsigned key = groups * tiles;
    if (s < n && i < cnt[s * m + rem / ns]) {
        const unsigned sp = min((unsigned)S[s * rowLen + rem], n - 1);
        key = ((tile / tilesPerPart) * chunks + sp / chunkStates) * tiles + tile;
    }
    keys[x] = key;
    vals[x] = (unsigned)x;
}

// segOff[k] = the first sorted index whose key is >= k, for k = 0 .. chunks * tiles
__global__ __launch_bounds__(256) void vi_chunk_bounds(size_t total, unsigned keyEnd,
        const unsigned *__restrict__ keys, unsigned *__restrict__ segOff)
{
    const size_t x = blockIdx.x * (size_t)256 + threadIdx.x;
    if (x >= total)
        return;
    const unsigned k = keys[x];
    const unsigned kp = x > 0 ? keys[x - 1] : 0u;
    const unsigned lo = x > 0 ? kp + 1 : 0u;
    for (unsigned kk = lo; kk <= k && kk <= keyEnd; kk++)
        segOff[kk] = (unsigned)x;
    if (x + 1 == total)
        for (unsigned kk = k + 1; kk <= keyEnd; kk++)
            segOff[kk] = (unsigned)total;
}

// the stored entry of every valid sorted index: successor offset in its chunk, T, slot in the
// tile (entry-major: i * rowsPerTile + local row), at its chunk-aligned position
__global__ __launch_bounds__(256) void vi_chunk_finish(unsigned valid, unsigned n, unsigned m,
        unsigned ns, unsigned statesPerTile, unsigned tileEntries, unsigned chunkStates,
        unsigned tiles, unsigned chunks, const unsigned *__restrict__ keys,
        const unsigned *__restrict__ vals, const unsigned *__restrict__ segOff,
        const unsigned *__restrict__ groupBase,
        const int *__restrict__ S, const float *__restrict__ T,
        unsigned short *__restrict__ succ, float *__restrict__ tOut,
        unsigned short *__restrict__ slot)
{
    const unsigned x = blockIdx.x * 256 + threadIdx.x;
    if (x >= valid)
        return;
    const unsigned key = keys[x], g = key / tiles, c = g % chunks;
    const unsigned pos = groupBase[g] + (x - segOff[g * tiles]);
    const unsigned v = vals[x];
    const unsigned tile = v / tileEntries, e = v % tileEntries;
    const unsigned rowLen = m * ns, rowsPerTile = statesPerTile * m;
    const size_t s = (size_t)tile * statesPerTile + e / rowLen;
    const unsigned rem = e % rowLen, i = rem % ns, a = rem / ns;
    const size_t flat = s * rowLen + rem;
    const unsigned sp = min((unsigned)S[flat], n - 1);
    succ[pos] = (unsigned short)(sp - c * chunkStates);
    tOut[pos] = T[flat];
    slot[pos] = (unsigned short)(i * rowsPerTile + (e / rowLen) * m + a);
}

// per tile, chunks + 1 entries: (stored position of the tile's segment of chunk c, entries of
// the tile in chunks before c); entry `chunks` holds the tile's total. One lane per tile.
__global__ __launch_bounds__(256) void vi_chunk_tile_segments(unsigned tiles, unsigned chunks,
        unsigned tilesPerPart, const unsigned *__restrict__ segOff,
        const unsigned *__restrict__ groupBase, uint2 *__restrict__ tileSeg)
{
    const unsigned t = blockIdx.x * 256 + threadIdx.x;
    if (t >= tiles)
        return;
    unsigned pre = 0;
    uint2 *row = tileSeg + (size_t)t * (chunks + 1);
    for (unsigned c = 0; c < chunks; c++) {
        const unsigned g = (t / tilesPerPart) * chunks + c;
        const unsigned k = g * tiles + t;
        row[c] = make_uint2(groupBase[g] + (segOff[k] - segOff[g * tiles]), pre);
        pre += segOff[k + 1] - segOff[k];
    }
    row[chunks] = make_uint2(0u, pre);
}

typedef unsigned short ushort8v __attribute__((ext_vector_type(8)));

// Pass A. A piece is a run of one chunk's stored entries, kChunkAlign-aligned, whose length is
// a multiple of kChunkThreads * kChunkVec entries or ends inside the chunk's zero padding
// (successor offset 0, T 0), so every lane handles whole groups of kChunkVec entries.
__global__ __launch_bounds__(kChunkThreads) void vi_chunk_products(unsigned n,
        unsigned chunkStates, float gamma, unsigned pieceBase, const uint4 *__restrict__ pieces,
        const unsigned short *__restrict__ succ, const float *__restrict__ T,
        const float *__restrict__ V, float *__restrict__ P)
{
    extern __shared__ float vs[];
    const uint4 pc = pieces[pieceBase + blockIdx.x];   // chunk, first entry, count (of 8s)
    const unsigned c = pc.x, t = threadIdx.x;
    const unsigned base = c * chunkStates;
    if (base >= n)
        return;               // a piece of another layout's chunking: touch nothing
    const unsigned cntV = min(chunkStates, n - base);
    // the chunk of V into LDS: 16-byte loads, eight per lane requested before any is stored
    // (chunk starts are multiples of 4,096 states, so V + base is 16-byte aligned)
    {
        const unsigned nv4 = cntV / 4;
        const float4v *V4 = reinterpret_cast<const float4v *>(V + base);
        float4v *vs4 = reinterpret_cast<float4v *>(vs);
        for (unsigned x0 = t; x0 < nv4; x0 += 8 * kChunkThreads) {
            float4v tmp[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const unsigned x = x0 + u * kChunkThreads;
                tmp[u] = V4[min(x, nv4 - 1)];
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const unsigned x = x0 + u * kChunkThreads;
                if (x < nv4)
                    vs4[x] = tmp[u];
            }
        }
        for (unsigned x = nv4 * 4 + t; x < cntV; x += kChunkThreads)
            vs[x] = V[base + x];
    }
    __syncthreads();
    const size_t first = pc.y;
    const unsigned groups = pc.z / kChunkVec;
    // kChunkRounds groups of 8 entries per lane in flight: all stream loads (indices clamped
    // into the piece), then the LDS gathers and products, then the stores of groups in range
    constexpr int kChunkRounds = 4;
    for (unsigned g0 = t; g0 < groups; g0 += kChunkRounds * kChunkThreads) {
        ushort8v sv[kChunkRounds];
        float4v ta[kChunkRounds], tb[kChunkRounds];
#pragma unroll
        for (int u = 0; u < kChunkRounds; u++) {
            const unsigned g = g0 + u * kChunkThreads;
            const size_t at = first + (size_t)min(g, groups - 1) * kChunkVec;
            sv[u] = __builtin_nontemporal_load(reinterpret_cast<const ushort8v *>(succ + at));
            ta[u] = __builtin_nontemporal_load(reinterpret_cast<const float4v *>(T + at));
            tb[u] = __builtin_nontemporal_load(reinterpret_cast<const float4v *>(T + at + 4));
        }
        float4v pa[kChunkRounds], pb[kChunkRounds];
#pragma unroll
        for (int u = 0; u < kChunkRounds; u++) {
            pa[u].x = __fmul_rn(__fmul_rn(gamma, ta[u].x), vs[sv[u][0]]);
            pa[u].y = __fmul_rn(__fmul_rn(gamma, ta[u].y), vs[sv[u][1]]);
            pa[u].z = __fmul_rn(__fmul_rn(gamma, ta[u].z), vs[sv[u][2]]);
            pa[u].w = __fmul_rn(__fmul_rn(gamma, ta[u].w), vs[sv[u][3]]);
            pb[u].x = __fmul_rn(__fmul_rn(gamma, tb[u].x), vs[sv[u][4]]);
            pb[u].y = __fmul_rn(__fmul_rn(gamma, tb[u].y), vs[sv[u][5]]);
            pb[u].z = __fmul_rn(__fmul_rn(gamma, tb[u].z), vs[sv[u][6]]);
            pb[u].w = __fmul_rn(__fmul_rn(gamma, tb[u].w), vs[sv[u][7]]);
        }
#pragma unroll
        for (int u = 0; u < kChunkRounds; u++) {
            const unsigned g = g0 + u * kChunkThreads;
            if (g < groups) {
                const size_t at = first + (size_t)g * kChunkVec;
                __builtin_nontemporal_store(pa[u], reinterpret_cast<float4v *>(P + at));
                __builtin_nontemporal_store(pb[u], reinterpret_cast<float4v *>(P + at + 4));
            }
        }
    }
}

// Pass B: one workgroup per state tile. The tile's entries are numbered flat across its
// chunk segments (a prefix table in LDS, and a table from each block of 64 flat positions to
// its first segment); each lane requests kRowLoads entries (product and slot) at once and
// places the products into LDS by slot. Then every row is summed in order from R and the max
// over actions taken.
constexpr unsigned kMaxChunks = 1024;   // chunks a layout may have (the prefix table in LDS)
constexpr size_t kMaxStored = (size_t)1 << 30;   // stored entries: byte offsets fit 32 bits
constexpr int kRowLoads = 16;           // a tile of 16,384 slots in one round of loads

// (at most 64 registers: 32 waves per CU, two workgroups of 1,024 lanes or four of 512)
template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(8)))
void vi_chunk_rows(unsigned n, unsigned m,
        unsigned statesPerTile, unsigned chunks, unsigned tileBase,
        const uint2 *__restrict__ tileSeg, const float *__restrict__ P,
        const unsigned short *__restrict__ slot, const unsigned char *__restrict__ cnt,
        const float *__restrict__ R, float *__restrict__ Vp, unsigned *__restrict__ pi,
        int writePi)
{
    extern __shared__ float lds[];
    __shared__ uint4 blockMap[kSortedEntries / 64];
    const unsigned rowsPerTile = statesPerTile * m;
    uint2 *seg = reinterpret_cast<uint2 *>(lds);     // chunks + 1 (begin, prefix) pairs
    unsigned char *cntL = reinterpret_cast<unsigned char *>(lds + 2 * (chunks + 1));
    float *Q = lds + 2 * (chunks + 1) + (rowsPerTile + 3) / 4;   // rowsPerTile row sums
    float *Pl = Q + rowsPerTile;                     // the tile's products by slot
    // workgroups are dealt to the 8 XCDs in turn; XCD x takes a contiguous run of this launch's
    // tiles, so neighbouring tiles, which share the cache lines at their segments' ends, are
    // read through the same L2
    const unsigned nb = gridDim.x, xcd = blockIdx.x % 8, q8 = nb / 8, rem8 = nb % 8;
    const unsigned tile = tileBase + xcd * q8 + min(xcd, rem8) + blockIdx.x / 8;
    const unsigned t = threadIdx.x;
    const unsigned s0 = tile * statesPerTile;
    const unsigned states = min(statesPerTile, n - s0);
    const unsigned rows = states * m;
    // the segment table, and R and the valid lengths of the rows (Q starts as R), requested
    // together so that their latencies overlap
    const size_t segBase = (size_t)tile * (chunks + 1), g0 = (size_t)s0 * m;
    uint2 sg0 = make_uint2(0u, 0u);
    float r0 = 0.0f;
    unsigned c0 = 0;
    if (t <= chunks)
        sg0 = tileSeg[segBase + t];
    if (t < rows) {
        r0 = R[g0 + t];
        c0 = cnt[g0 + t];
    }
    if (t <= chunks)
        seg[t] = sg0;
    if (t < rows) {
        Q[t] = r0;
        cntL[t] = (unsigned char)c0;
    }
    for (unsigned c = t + NT; c <= chunks; c += NT)
        seg[c] = tileSeg[segBase + c];
    for (unsigned row = t + NT; row < rows; row += NT) {
        Q[row] = R[g0 + row];
        cntL[row] = cnt[g0 + row];
    }
    __syncthreads();
    // block b of 64 flat positions: the segment c holding position 64 b (ceil(prefix / 64) is
    // the first block starting inside a segment) and the one after it, as (stored position
    // minus flat position, end) pairs. A position of the block finds its stored position with
    // one table read when the block meets at most those two segments (segments are hundreds
    // of entries long), else by a binary search over the prefix table.
    for (unsigned c = t; c < chunks; c += NT) {
        const unsigned b1 = (seg[c + 1].y + 63) / 64;
        uint4 e;
        e.x = seg[c].x - seg[c].y;
        e.y = seg[c + 1].y;
        e.z = c + 1 < chunks ? seg[c + 1].x - seg[c + 1].y : 0u;
        e.w = c + 1 < chunks ? seg[c + 2].y : seg[c + 1].y;
        for (unsigned b = (seg[c].y + 63) / 64; b < b1; b++)
            blockMap[b] = e;
    }
    __syncthreads();
    const unsigned total = seg[chunks].y;
    const __amdgpu_buffer_rsrc_t rP = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>(P), 0, -1, kRawBufferWord3);
    const __amdgpu_buffer_rsrc_t rS = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<unsigned short *>(slot), 0, -1, kRawBufferWord3);
    for (unsigned f0 = t; f0 < total; f0 += kRowLoads * NT) {
        float p[kRowLoads];
        unsigned short sl[kRowLoads];
        unsigned at[kRowLoads];
        unsigned far = 0;                               // positions past the block's two
#pragma unroll
        for (int u = 0; u < kRowLoads; u++) {
            const unsigned f = min(f0 + u * NT, total - 1);
            const uint4 e = blockMap[f / 64];
            at[u] = f + (f < e.y ? e.x : e.z);
            far |= (f >= e.w ? 1u : 0u) << u;
        }
        if (far != 0) {                                 // a block meeting three segments
#pragma unroll
            for (int u = 0; u < kRowLoads; u++) {
                if ((far >> u) & 1u) {
                    const unsigned f = min(f0 + u * NT, total - 1);
                    unsigned lo = 0, hi = chunks;       // seg[lo].y <= f < seg[hi].y
                    while (hi - lo > 1) {
                        const unsigned mid = (lo + hi) / 2;
                        if (seg[mid].y <= f)
                            lo = mid;
                        else
                            hi = mid;
                    }
                    at[u] = seg[lo].x + (f - seg[lo].y);
                }
            }
        }
        // buffer loads take the 32-bit byte offset in one register (the layout keeps fewer
        // than 2^30 stored entries), streaming past the caches
#pragma unroll
        for (int u = 0; u < kRowLoads; u++) {
            p[u] = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(rP, at[u] * 4u, 0, 2));
            sl[u] = __builtin_amdgcn_raw_buffer_load_b16(rS, at[u] * 2u, 0, 2);
        }
#pragma unroll
        for (int u = 0; u < kRowLoads; u++)
            if (f0 + u * NT < total)
                Pl[sl[u]] = p[u];
    }
    __syncthreads();
    for (unsigned row = t; row < rows; row += NT) {
        float q = Q[row];
        const unsigned c = cntL[row];
        for (unsigned i = 0; i < c; i++)
            q = __fadd_rn(q, Pl[i * rowsPerTile + row]);
        Q[row] = q;
    }
    __syncthreads();
    for (unsigned x = t; x < states; x += NT) {
        const float *qs = Q + (size_t)x * m;
        float best = qs[0];
        unsigned bestA = 0;
        for (unsigned a = 1; a < m; a++) {
            if (qs[a] > best) {
                best = qs[a];
                bestA = a;
            }
        }
        Vp[s0 + x] = best;
        if (writePi)
            pi[s0 + x] = bestA;
    }
}

// ---- batched value iteration: many MDPs of uneven size, one launch per sweep for all
//
// Every sweep of every instance is captured into one HIP graph: launch k runs sweep k of all
// instances whose horizon exceeds k, over one grid holding every instance's tiles (an
// instance's workgroups are consecutive; its tile offset comes from a prefix table). A tile
// is kBatchRows (s, a) rows of whole states. Lane t of the 256 sums rows t, t + 256, ... in
// the reference's order, (gamma*T)*V products and sums rounded separately, and the lane of
// each state's action 0 takes the max over actions with the reference's tie rule: V' and pi
// equal the plain sweep's bit for bit.
struct ViInst {
    unsigned n, m, ns, horizon, statesPerTile, tiles;
    float gamma;
    const int *S;
    const float *T, *R;
    float *V, *Vp;
    unsigned *pi;
};

constexpr int kBatchThreads = 256;
constexpr unsigned kBatchRows = 1024;

__global__ __launch_bounds__(kBatchThreads) void vi_batch_sweep(const ViInst *__restrict__ inst,
        const unsigned *__restrict__ off, unsigned count, unsigned k)
{
    __shared__ float q[kBatchRows];
    const unsigned id = blockIdx.x;
    unsigned lo = 0, hi = count;               // off[lo] <= id < off[hi]
    while (hi - lo > 1) {
        const unsigned mid = (lo + hi) / 2;
        if (off[mid] <= id)
            lo = mid;
        else
            hi = mid;
    }
    const ViInst I = inst[lo];
    if (k >= I.horizon)
        return;
    const bool even = (k % 2) == 0;
    const float *V = even ? I.V : I.Vp;
    float *Vp = even ? I.Vp : I.V;
    const unsigned tile = id - off[lo];
    const unsigned s0 = tile * I.statesPerTile;
    const unsigned states = min(I.statesPerTile, I.n - s0);
    const unsigned rows = states * I.m;
    const size_t rowBase = (size_t)s0 * I.m;
    for (unsigned r = threadIdx.x; r < rows; r += kBatchThreads) {
        const size_t row = rowBase + r;
        const int *Sr = I.S + row * I.ns;
        const float *Tr = I.T + row * I.ns;
        float acc = I.R[row];
        for (unsigned i = 0; i < I.ns; i++) {
            const int sp = Sr[i];
            if (sp < 0)
                break;
            acc = __fadd_rn(acc, __fmul_rn(__fmul_rn(I.gamma, Tr[i]),
                                           V[min((unsigned)sp, I.n - 1)]));
        }
        q[r] = acc;
    }
    __syncthreads();
    const bool last = k + 1 == I.horizon;
    for (unsigned x = threadIdx.x; x < states; x += kBatchThreads) {
        const float *qs = q + (size_t)x * I.m;
        float best = qs[0];
        unsigned bestA = 0;
        for (unsigned a = 1; a < I.m; a++) {
            if (qs[a] > best) {
                best = qs[a];
                bestA = a;
            }
        }
        Vp[s0 + x] = best;
        if (last)
            I.pi[s0 + x] = bestA;
    }
}

unsigned pow2_at_least(unsigned x)
{
    unsigned p = 1;
    while (p < x)
        p <<= 1;
    return p;
}

} // namespace

// Every sweep layout built gets a serial number of its own. The cached sweep graph is keyed by
// it (not by the layout's address, which a later layout of other settings may reuse) and is
// dropped with the layout it was captured on.
static unsigned long long g_layoutSerial = 0;
static void vi_graph_forget_layout(unsigned long long serial);

// The successor-sorted layout of one solver's model (built by mdp_vi_initialize_gpu for large
// models, freed with the solver; keyed by the solver's V buffer and the model's arrays).
struct ViSorted {
    unsigned long long serial = 0;
    const float *owner = nullptr;
    const int *dS = nullptr;
    const float *dT = nullptr;
    unsigned long long generation = 0;   // model_generation(true) when built
    unsigned n = 0, m = 0, ns = 0, statesPerTile = 0, tileEntries = 0, tiles = 0;
    void *mem = nullptr;
    int *sS = nullptr;
    float *sT = nullptr;
    unsigned short *slot = nullptr;
    unsigned char *cnt = nullptr;
    unsigned *tileValid = nullptr;
    size_t ldsBytes = 0;
};
static std::vector<ViSorted> g_sorted;

// The layout of this solver and model, unless S or T were uploaded or released through this
// library since it was built (it holds copies of them): sweeps then use the wave-tiled kernel,
// which reads d_S and d_T, until mdp_vi_initialize_gpu builds the layout again.
static const ViSorted *sorted_for(const nova::MDP *mdp, const nova::MDPVIGPU *vi)
{
    for (const ViSorted &L : g_sorted)
        if (L.owner == vi->d_V && L.dS == mdp->d_S && L.dT == mdp->d_T && L.n == mdp->n &&
                L.m == mdp->m && L.ns == mdp->ns && L.generation == model_generation(true))
            return &L;
    return nullptr;
}

static void sorted_drop(const float *owner, const char *fn)
{
    for (size_t x = 0; x < g_sorted.size();) {
        if (g_sorted[x].owner == owner) {
            hipDeviceSynchronize();
            vi_graph_forget_layout(g_sorted[x].serial);
            device_free(g_sorted[x].mem, fn, "the sorted sweep layout");
            g_sorted.erase(g_sorted.begin() + (long)x);
        } else {
            x++;
        }
    }
}

// Which layout initialize builds: 0 none (the wave-tiled sweep reads d_S and d_T), 2 the
// successor-sorted tiles, 3 the successor chunks. Both layouts take ns <= 255 (byte row
// counts) and m*ns <= kSortedEntries (16-bit slots, LDS for a tile's products); the sorted
// one also needs ns >= 8. NOVA_AMD_VI_KERNEL=tiled|sorted|chunked forces a choice where the
// shape allows it; by default models of at least 2^24 (s, a, i) slots, where the wave-tiled
// sweep is bound by gather requests, get the chunked layout.
static int layout_wanted(const nova::MDP *mdp)
{
    const size_t slots = (size_t)mdp->n * mdp->m * mdp->ns;
    const bool common = mdp->ns <= 255 && (size_t)mdp->m * mdp->ns <= kSortedEntries &&
                        slots + kSortedEntries < ((size_t)1 << 31);
    if (!common)
        return 0;
    const bool sortedOk = mdp->ns >= 8;
    // (the chunk groups' alignment padding stays far below the margin of 2^26 entries)
    const bool chunkedOk = (size_t)mdp->n <= (size_t)kMaxChunks * 32768 &&
                           slots + ((size_t)1 << 26) < kMaxStored;
    if (const char *f = std::getenv("NOVA_AMD_VI_KERNEL")) {
        if (std::strcmp(f, "tiled") == 0)
            return 0;
        if (std::strcmp(f, "sorted") == 0)
            return sortedOk ? 2 : 0;
        if (std::strcmp(f, "chunked") == 0)
            return chunkedOk ? 3 : 0;
    }
    if (slots < ((size_t)1 << 24))
        return 0;
    return chunkedOk ? 3 : (sortedOk ? 2 : 0);
}

static size_t up256(size_t x) { return (x + 255) / 256 * 256; }

static int sorted_build(const nova::MDP *mdp, const float *owner, hipStream_t st, const char *fn)
{
    const unsigned n = mdp->n, m = mdp->m, ns = mdp->ns;
    ViSorted L;
    L.owner = owner;
    L.generation = model_generation(true);
    L.dS = mdp->d_S;
    L.dT = mdp->d_T;
    L.n = n;
    L.m = m;
    L.ns = ns;
    L.statesPerTile = kSortedEntries / (m * ns);
    L.tileEntries = L.statesPerTile * m * ns;
    L.tiles = (n + L.statesPerTile - 1) / L.statesPerTile;
    L.ldsBytes = sizeof(float) * ((size_t)L.tileEntries + (size_t)L.statesPerTile * m);
    const size_t total = (size_t)L.tiles * L.tileEntries, rows = (size_t)n * m;
    const size_t bS = up256(4 * total), bT = up256(4 * total), bSl = up256(2 * total);
    const size_t bC = up256(rows), bV = up256(4 * (size_t)L.tiles);
    int rc = device_alloc(&L.mem, bS + bT + bSl + bC + bV, fn, "the sorted sweep layout");
    if (rc != NOVA_SUCCESS)
        return rc;
    char *c = static_cast<char *>(L.mem);
    L.sS = reinterpret_cast<int *>(c);
    L.sT = reinterpret_cast<float *>(c + bS);
    L.slot = reinterpret_cast<unsigned short *>(c + bS + bT);
    L.cnt = reinterpret_cast<unsigned char *>(c + bS + bT + bSl);
    L.tileValid = reinterpret_cast<unsigned *>(c + bS + bT + bSl + bC);

    // temporaries: keys, slot values in and out, segment offsets, sort storage
    void *tmp = nullptr, *cub = nullptr;
    const size_t bK = up256(4 * total), bO = up256(4 * ((size_t)L.tiles + 1));
    rc = device_alloc(&tmp, bK + 2 * bK + bO, fn, "sort buffers");
    if (rc != NOVA_SUCCESS) {
        device_free(L.mem, fn, "the sorted sweep layout");
        return rc;
    }
    char *tc = static_cast<char *>(tmp);
    int *keys = reinterpret_cast<int *>(tc);
    unsigned *valsIn = reinterpret_cast<unsigned *>(tc + bK);
    unsigned *valsOut = reinterpret_cast<unsigned *>(tc + 2 * bK);
    int *off = reinterpret_cast<int *>(tc + 3 * bK);
    std::vector<int> hoff((size_t)L.tiles + 1);
    for (size_t x = 0; x <= L.tiles; x++)
        hoff[x] = (int)(x * L.tileEntries);
    hipError_t e = hipMemcpy(off, hoff.data(), sizeof(int) * hoff.size(), hipMemcpyHostToDevice);
    const unsigned gR = (unsigned)((rows + 255) / 256), gE = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(vi_row_counts, dim3(gR), dim3(256), 0, st, rows, ns, mdp->d_S, L.cnt);
    hipLaunchKernelGGL(vi_sort_keys, dim3(gE), dim3(256), 0, st, total, n, m, ns,
                       L.statesPerTile, L.tileEntries, mdp->d_S, L.cnt, keys, valsIn);
    int endBit = 1;
    while (endBit < 31 && ((size_t)1 << endBit) <= n)
        endBit++;
    size_t cubBytes = 0;
    if (e == hipSuccess)
        e = hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, cubBytes, keys, L.sS, valsIn,
                valsOut, (int)total, (int)L.tiles, off, off + 1, 0, endBit, st);
    if (e == hipSuccess)
        rc = device_alloc(&cub, std::max<size_t>(cubBytes, 256), fn, "sort storage");
    if (e == hipSuccess && rc == NOVA_SUCCESS) {
        e = hipcub::DeviceSegmentedRadixSort::SortPairs(cub, cubBytes, keys, L.sS, valsIn,
                valsOut, (int)total, (int)L.tiles, off, off + 1, 0, endBit, st);
        if (e == hipSuccess)
            e = hipMemsetAsync(L.tileValid, 0, 4 * (size_t)L.tiles, st);
        hipLaunchKernelGGL(vi_sort_finish, dim3(gE), dim3(256), 0, st, total, n, m, ns,
                           L.statesPerTile, L.tileEntries, mdp->d_T, L.sS, valsOut, L.sT, L.slot,
                           L.tileValid);
        if (e == hipSuccess)
            e = hipGetLastError();
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
    }
    if (cub)
        device_free(cub, fn, "sort storage");
    device_free(tmp, fn, "sort buffers");
    if (e == hipSuccess && rc == NOVA_SUCCESS)
        e = hipFuncSetAttribute((const void *)vi_sweep_sorted,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(float) * 2 * (size_t)kSortedEntries));
    if (e != hipSuccess || rc != NOVA_SUCCESS) {
        device_free(L.mem, fn, "the sorted sweep layout");
        return rc != NOVA_SUCCESS ? rc : hip_fail(e, fn, "Failed to build the sorted layout.",
                                                  NOVA_ERROR_KERNEL_EXECUTION);
    }
    L.serial = ++g_layoutSerial;
    g_sorted.push_back(L);
    return NOVA_SUCCESS;
}

// The successor-chunked layout of one solver's model (built by mdp_vi_initialize_gpu for large
// models, freed with the solver; keyed like the sorted layout).
struct ViChunked {
    unsigned long long serial = 0;
    const float *owner = nullptr;
    const int *dS = nullptr;
    const float *dT = nullptr;
    unsigned long long generation = 0;
    unsigned n = 0, m = 0, ns = 0, statesPerTile = 0, tileEntries = 0, tiles = 0;
    unsigned chunkStates = 0, chunks = 0, pieceCount = 0;
    unsigned parts = 1, tilesPerPart = 0;
    std::vector<unsigned> partPieces;   // parts + 1 prefix offsets into the pieces
    void *mem = nullptr;
    unsigned short *succ = nullptr;
    float *T = nullptr;
    unsigned short *slot = nullptr;
    float *P = nullptr;
    unsigned char *cnt = nullptr;
    uint2 *tileSeg = nullptr;
    uint4 *pieces = nullptr;
    size_t ldsA = 0, ldsB = 0;
};
static std::vector<ViChunked> g_chunked;

static const ViChunked *chunked_for(const nova::MDP *mdp, const nova::MDPVIGPU *vi)
{
    for (const ViChunked &L : g_chunked)
        if (L.owner == vi->d_V && L.dS == mdp->d_S && L.dT == mdp->d_T && L.n == mdp->n &&
                L.m == mdp->m && L.ns == mdp->ns && L.generation == model_generation(true))
            return &L;
    return nullptr;
}

static void chunked_drop(const float *owner, const char *fn)
{
    for (size_t x = 0; x < g_chunked.size();) {
        if (g_chunked[x].owner == owner) {
            hipDeviceSynchronize();
            vi_graph_forget_layout(g_chunked[x].serial);
            device_free(g_chunked[x].mem, fn, "the chunked sweep layout");
            g_chunked.erase(g_chunked.begin() + (long)x);
        } else {
            x++;
        }
    }
}

static unsigned env_unsigned(const char *name, unsigned dflt)
{
    if (const char *v = std::getenv(name)) {
        const long x = std::strtol(v, nullptr, 10);
        if (x > 0)
            return (unsigned)x;
    }
    return dflt;
}

static int chunked_build(const nova::MDP *mdp, const float *owner, hipStream_t st,
                         const char *fn)
{
    const unsigned n = mdp->n, m = mdp->m, ns = mdp->ns;
    ViChunked L;
    L.owner = owner;
    L.generation = model_generation(true);
    L.dS = mdp->d_S;
    L.dT = mdp->d_T;
    L.n = n;
    L.m = m;
    L.ns = ns;
    // tiles of 8,192 slots (NOVA_AMD_VI_TILE: 4,096 to 16,384), whole states each: pass B then
    // runs four 512-lane workgroups per CU rather than two of 1,024 lanes
    const unsigned tileSlots = std::min(kSortedEntries, std::max(std::max(4096u, m * ns),
            env_unsigned("NOVA_AMD_VI_TILE", kChunkTileDefault)));
    L.statesPerTile = tileSlots / (m * ns);
    L.tileEntries = L.statesPerTile * m * ns;
    L.tiles = (n + L.statesPerTile - 1) / L.statesPerTile;
    // chunk of successor states whose V fills a pass-A workgroup's LDS (a power of two from
    // 4,096 to 32,768; NOVA_AMD_VI_CHUNK overrides), pieces of whole load rounds
    unsigned cs = env_unsigned("NOVA_AMD_VI_CHUNK", kChunkStatesDefault);
    cs = std::min(32768u, std::max(4096u, cs / 4096 * 4096));   // 16-byte aligned chunk starts
    while ((n + cs - 1) / cs > kMaxChunks && cs < 32768)
        cs *= 2;                       // pass B keeps a prefix entry per chunk in LDS
    L.chunkStates = cs;
    L.chunks = (n + cs - 1) / cs;
    const unsigned round = kChunkThreads * kChunkVec;
    unsigned piece = env_unsigned("NOVA_AMD_VI_PIECE", kPieceEntriesDefault);
    piece = std::max(round, piece / round * round);
    const size_t total = (size_t)L.tiles * L.tileEntries, rows = (size_t)n * m;
    // row parts: each sweep runs pass A and pass B part by part, so that one part's products
    // (1 / parts of them) can stay in the last-level cache between the passes
    unsigned parts = env_unsigned("NOVA_AMD_VI_PARTS", 1);
    parts = std::min(parts, L.tiles);
    L.tilesPerPart = (L.tiles + parts - 1) / parts;
    L.parts = (L.tiles + L.tilesPerPart - 1) / L.tilesPerPart;
    const unsigned groups = L.parts * L.chunks;
    const unsigned keyEnd = groups * L.tiles;

    // temporaries: row counts (kept), keys and values in and out, segment offsets
    void *tmp = nullptr, *cub = nullptr;
    const size_t bK = up256(4 * total), bSeg = up256(4 * ((size_t)keyEnd + 1));
    const size_t bC = up256(rows);
    int rc = device_alloc(&tmp, 4 * bK + bSeg + bC, fn, "layout build buffers");
    if (rc != NOVA_SUCCESS)
        return rc;
    char *tc = static_cast<char *>(tmp);
    unsigned *keysIn = reinterpret_cast<unsigned *>(tc);
    unsigned *keysOut = reinterpret_cast<unsigned *>(tc + bK);
    unsigned *valsIn = reinterpret_cast<unsigned *>(tc + 2 * bK);
    unsigned *valsOut = reinterpret_cast<unsigned *>(tc + 3 * bK);
    unsigned *segOff = reinterpret_cast<unsigned *>(tc + 4 * bK);
    unsigned char *cntTmp = reinterpret_cast<unsigned char *>(tc + 4 * bK + bSeg);
    const unsigned gR = (unsigned)((rows + 255) / 256), gE = (unsigned)((total + 255) / 256);
    hipLaunchKernelGGL(vi_row_counts, dim3(gR), dim3(256), 0, st, rows, ns, mdp->d_S, cntTmp);
    hipLaunchKernelGGL(vi_chunk_keys, dim3(gE), dim3(256), 0, st, total, n, m, ns,
                       L.statesPerTile, L.tileEntries, cs, L.tiles, L.chunks, L.tilesPerPart,
                       groups, mdp->d_S, cntTmp, keysIn, valsIn);
    int endBit = 1;
    while (endBit < 32 && ((size_t)1 << endBit) <= keyEnd)
        endBit++;
    size_t cubBytes = 0;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, cubBytes, keysIn, keysOut,
                                                      valsIn, valsOut, (int)total, 0, endBit, st);
    if (e == hipSuccess)
        rc = device_alloc(&cub, std::max<size_t>(cubBytes, 256), fn, "sort storage");
    if (e == hipSuccess && rc == NOVA_SUCCESS)
        e = hipcub::DeviceRadixSort::SortPairs(cub, cubBytes, keysIn, keysOut, valsIn, valsOut,
                                               (int)total, 0, endBit, st);
    if (e == hipSuccess)
        hipLaunchKernelGGL(vi_chunk_bounds, dim3(gE), dim3(256), 0, st, total, keyEnd, keysOut,
                           segOff);
    // group boundaries on the host: group g = (part, chunk) starts at a kChunkAlign-aligned
    // position, and its zero padding covers the rounding of its last piece up to whole load
    // rounds of 8; the pieces are listed part by part
    std::vector<unsigned> seg((size_t)keyEnd + 1);
    if (e == hipSuccess)
        e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(seg.data(), segOff, 4 * seg.size(), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess)
        e = hipStreamSynchronize(st);
    std::vector<unsigned> chunkBase((size_t)groups + 1);   // the groups' stored positions
    std::vector<uint4> pieces;
    L.partPieces.assign((size_t)L.parts + 1, 0);
    size_t stored = 0;
    if (e == hipSuccess && rc == NOVA_SUCCESS) {
        for (unsigned g = 0; g < groups; g++) {
            const unsigned c = g % L.chunks;
            chunkBase[g] = (unsigned)stored;
            const unsigned len = seg[(size_t)(g + 1) * L.tiles] - seg[(size_t)g * L.tiles];
            for (unsigned off = 0; off < len; off += piece) {
                const unsigned cntE = std::min(piece, len - off);
                pieces.push_back(make_uint4(c, (unsigned)stored + off,
                                            (cntE + kChunkVec - 1) / kChunkVec * kChunkVec, 0));
            }
            stored += (len + kChunkAlign - 1) / kChunkAlign * kChunkAlign;
            if (c + 1 == L.chunks)
                L.partPieces[g / L.chunks + 1] = (unsigned)pieces.size();
        }
        chunkBase[groups] = (unsigned)stored;
        if (stored >= kMaxStored || pieces.empty())
            rc = NOVA_ERROR_INVALID_DATA;
    }
    L.pieceCount = (unsigned)pieces.size();
    // the layout: successor offsets, T, slots, products (stored entries each), row counts,
    // tile segments, pieces, chunk bases
    const size_t bSu = up256(2 * stored), bT = up256(4 * stored), bSl = up256(2 * stored);
    const size_t bP = up256(4 * stored), bTS = up256(8 * (size_t)L.tiles * (L.chunks + 1));
    const size_t bPc = up256(16 * pieces.size()), bCB = up256(4 * chunkBase.size());
    if (e == hipSuccess && rc == NOVA_SUCCESS)
        rc = device_alloc(&L.mem, bSu + bT + bSl + bP + bC + bTS + bPc + bCB, fn,
                          "the chunked sweep layout");
    if (e == hipSuccess && rc == NOVA_SUCCESS) {
        char *c = static_cast<char *>(L.mem);
        L.succ = reinterpret_cast<unsigned short *>(c);
        L.T = reinterpret_cast<float *>(c + bSu);
        L.slot = reinterpret_cast<unsigned short *>(c + bSu + bT);
        L.P = reinterpret_cast<float *>(c + bSu + bT + bSl);
        L.cnt = reinterpret_cast<unsigned char *>(c + bSu + bT + bSl + bP);
        L.tileSeg = reinterpret_cast<uint2 *>(c + bSu + bT + bSl + bP + bC);
        L.pieces = reinterpret_cast<uint4 *>(c + bSu + bT + bSl + bP + bC + bTS);
        unsigned *dBase = reinterpret_cast<unsigned *>(c + bSu + bT + bSl + bP + bC + bTS + bPc);
        // zero padding: offset 0 and T 0 (a finite product no row reads)
        e = hipMemsetAsync(L.succ, 0, bSu + bT, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(L.cnt, cntTmp, rows, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(L.pieces, pieces.data(), 16 * pieces.size(),
                               hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dBase, chunkBase.data(), 4 * chunkBase.size(),
                               hipMemcpyHostToDevice, st);
        const unsigned valid = seg[keyEnd];
        if (e == hipSuccess && valid > 0)
            hipLaunchKernelGGL(vi_chunk_finish, dim3((valid + 255) / 256), dim3(256), 0, st,
                               valid, n, m, ns, L.statesPerTile, L.tileEntries, cs, L.tiles,
                               L.chunks, keysOut, valsOut, segOff, dBase, mdp->d_S, mdp->d_T,
                               L.succ, L.T, L.slot);
        if (e == hipSuccess)
            hipLaunchKernelGGL(vi_chunk_tile_segments, dim3((L.tiles + 255) / 256), dim3(256),
                               0, st, L.tiles, L.chunks, L.tilesPerPart, segOff, dBase,
                               L.tileSeg);
        if (e == hipSuccess)
            e = hipGetLastError();
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
    }
    if (cub)
        device_free(cub, fn, "sort storage");
    device_free(tmp, fn, "layout build buffers");
    L.ldsA = sizeof(float) * (size_t)cs;
    L.ldsB = sizeof(uint2) * ((size_t)L.chunks + 1) +
             ((size_t)L.statesPerTile * m + 3) / 4 * 4 +
             sizeof(float) * ((size_t)L.statesPerTile * m + L.tileEntries);
    // the largest dynamic LDS any layout launches with (chunks of 32,768 states; the segment
    // table of kMaxChunks chunks, row lengths, row sums and products of 16,384 slots each), so
    // that the graphs of layouts alive at once all launch
    if (e == hipSuccess && rc == NOVA_SUCCESS)
        e = hipFuncSetAttribute((const void *)vi_chunk_products,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(float) * 32768));
    if (e == hipSuccess && rc == NOVA_SUCCESS)
        e = hipFuncSetAttribute((const void *)vi_chunk_rows<1024>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(uint2) * (kMaxChunks + 1) + kSortedEntries +
                                      sizeof(float) * 2 * (size_t)kSortedEntries));
    if (e == hipSuccess && rc == NOVA_SUCCESS)
        e = hipFuncSetAttribute((const void *)vi_chunk_rows<512>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(sizeof(uint2) * (kMaxChunks + 1) + kSortedEntries +
                                      sizeof(float) * 2 * (size_t)kSortedEntries));
    if (e != hipSuccess || rc != NOVA_SUCCESS) {
        if (L.mem)
            device_free(L.mem, fn, "the chunked sweep layout");
        return rc != NOVA_SUCCESS ? rc : hip_fail(e, fn, "Failed to build the chunked layout.",
                                                  NOVA_ERROR_KERNEL_EXECUTION);
    }
    L.serial = ++g_layoutSerial;
    g_chunked.push_back(L);
    return NOVA_SUCCESS;
}

// A sweep launch plan for one MDP shape.
struct ViPlan {
    TiledFn tiled = nullptr;
    const ViSorted *sorted = nullptr;
    const ViChunked *chunked = nullptr;
    unsigned statesPerTile = 0;
    unsigned grid = 0;
};

static ViPlan vi_plan(const nova::MDP *mdp, const nova::MDPVIGPU *vi)
{
    ViPlan p;
    if ((p.chunked = chunked_for(mdp, vi)) != nullptr) {
        p.grid = p.chunked->tiles;
        return p;
    }
    if ((p.sorted = sorted_for(mdp, vi)) != nullptr) {
        p.grid = p.sorted->tiles;
        return p;
    }
    const unsigned n = mdp->n, m = mdp->m, ns = mdp->ns;
    const bool aligned = ((reinterpret_cast<uintptr_t>(mdp->d_S) |
                           reinterpret_cast<uintptr_t>(mdp->d_T)) & 15u) == 0;
    if (m <= (unsigned)kTileRows) {
        if (ns % 4 == 0 && aligned && ns <= 256)
            p.tiled = tiled_for<4>(pow2_at_least(ns / 4));
        else if (ns <= 64)
            p.tiled = tiled_for<1>(pow2_at_least(ns));
    }
    if (p.tiled) {
        p.statesPerTile = kTileRows / m;
        p.grid = (n + p.statesPerTile - 1) / p.statesPerTile;
    } else {
        p.grid = (n + kBlock - 1) / kBlock;
    }
    return p;
}

// writePi: the policy buffer is shared by all sweeps and only the last sweep's greedy actions
// are ever read (src/mdp/algorithms/mdp_vi_cpu.cpp:187-191, :215-220), so a run of sweeps
// writes pi on its final sweep only.
static void vi_launch(const ViPlan &p, const nova::MDP *mdp, const float *V, float *Vp,
                      unsigned *pi, bool writePi, hipStream_t st)
{
    if (p.chunked) {
        const ViChunked &L = *p.chunked;
        for (unsigned h = 0; h < L.parts; h++) {
            const unsigned p0 = L.partPieces[h], p1 = L.partPieces[h + 1];
            const unsigned t0 = h * L.tilesPerPart, t1 = std::min(L.tiles, t0 + L.tilesPerPart);
            if (p1 > p0)
                hipLaunchKernelGGL(vi_chunk_products, dim3(p1 - p0), dim3(kChunkThreads),
                                   L.ldsA, st, mdp->n, L.chunkStates, mdp->gamma, p0, L.pieces,
                                   L.succ, L.T, V, L.P);
            if (L.tileEntries <= kSortedEntries / 2)   // 512 lanes of 16 slots each
                hipLaunchKernelGGL(vi_chunk_rows<512>, dim3(t1 - t0), dim3(512), L.ldsB, st,
                                   mdp->n, mdp->m, L.statesPerTile, L.chunks, t0, L.tileSeg,
                                   L.P, L.slot, L.cnt, mdp->d_R, Vp, pi, (int)writePi);
            else
                hipLaunchKernelGGL(vi_chunk_rows<1024>, dim3(t1 - t0), dim3(1024), L.ldsB, st,
                                   mdp->n, mdp->m, L.statesPerTile, L.chunks, t0, L.tileSeg,
                                   L.P, L.slot, L.cnt, mdp->d_R, Vp, pi, (int)writePi);
        }
        return;
    }
    if (p.sorted) {
        const ViSorted &L = *p.sorted;
        hipLaunchKernelGGL(vi_sweep_sorted, dim3(p.grid), dim3(kSortedThreads), L.ldsBytes, st,
                           mdp->n, mdp->m, mdp->ns, L.statesPerTile, L.tileEntries, mdp->gamma,
                           L.sS, L.sT, L.slot, L.tileValid, L.cnt, mdp->d_R, V, Vp, pi,
                           (int)writePi);
        return;
    }
    if (p.tiled)
        hipLaunchKernelGGL(p.tiled, dim3(p.grid), dim3(kBlock), 0, st, mdp->n, mdp->m, mdp->ns,
                           p.statesPerTile, mdp->gamma, mdp->d_S, mdp->d_T, mdp->d_R, V, Vp, pi,
                           (int)writePi);
    else
        hipLaunchKernelGGL(vi_sweep_plain, dim3(p.grid), dim3(kBlock), 0, st, mdp->n, mdp->m,
                           mdp->ns, mdp->gamma, mdp->d_S, mdp->d_T, mdp->d_R, V, Vp, pi,
                           (int)writePi);
}

// Cached executable graph of a run of sweeps (one cache entry; re-captured when any launch
// argument changes).
struct ViGraph {
    hipGraphExec_t exec = nullptr;
    const void *key[8] = {};
    unsigned sweeps = 0, parity = 0, n = 0, m = 0, ns = 0;
    float gamma = 0.0f;
};
static ViGraph g_graph;

static void vi_graph_forget_layout(unsigned long long serial)
{
    if (g_graph.exec != nullptr && g_graph.key[7] == (const void *)(uintptr_t)serial) {
        hipGraphExecDestroy(g_graph.exec);
        g_graph.exec = nullptr;
    }
}

static int vi_run_sweeps(const nova::MDP *mdp, nova::MDPVIGPU *vi, unsigned sweeps,
                         hipStream_t st, hipEvent_t t0, hipEvent_t t1, const char *fn)
{
    const ViPlan plan = vi_plan(mdp, vi);
    const unsigned parity = vi->currentHorizon % 2;
    const unsigned long long layout = plan.chunked ? plan.chunked->serial
                                                   : (plan.sorted ? plan.sorted->serial : 0ull);
    const void *key[8] = {mdp->d_S, mdp->d_T, mdp->d_R, vi->d_V, vi->d_VPrime, vi->d_pi,
                          (const void *)plan.tiled, (const void *)(uintptr_t)layout};
    bool hit = g_graph.exec != nullptr && g_graph.sweeps == sweeps && g_graph.parity == parity &&
               g_graph.n == mdp->n && g_graph.m == mdp->m && g_graph.ns == mdp->ns &&
               g_graph.gamma == mdp->gamma && std::memcmp(g_graph.key, key, sizeof(key)) == 0;
    if (!hit) {
        if (g_graph.exec) {
            hipGraphExecDestroy(g_graph.exec);
            g_graph.exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        hipStream_t cs = capture_stream();
        if (cs == nullptr) {
            NOVA_ERR(fn, "Could not create the graph-capture stream.");
            return NOVA_ERROR_KERNEL_EXECUTION;
        }
        hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess)
            return hip_fail(e, fn, "Failed to begin graph capture.", NOVA_ERROR_KERNEL_EXECUTION);
        for (unsigned k = 0; k < sweeps; k++) {
            const bool even = ((parity + k) % 2) == 0;
            vi_launch(plan, mdp, even ? vi->d_V : vi->d_VPrime, even ? vi->d_VPrime : vi->d_V,
                      vi->d_pi, k + 1 == sweeps, cs);
        }
        e = hipStreamEndCapture(cs, &graph);
        if (e != hipSuccess)
            return hip_fail(e, fn, "Failed to capture the sweeps.", NOVA_ERROR_KERNEL_EXECUTION);
        e = hipGraphInstantiate(&g_graph.exec, graph, nullptr, nullptr, 0);
        hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g_graph.exec = nullptr;
            return hip_fail(e, fn, "Failed to instantiate the sweep graph.",
                            NOVA_ERROR_KERNEL_EXECUTION);
        }
        std::memcpy(g_graph.key, key, sizeof(key));
        g_graph.sweeps = sweeps;
        g_graph.parity = parity;
        g_graph.n = mdp->n;
        g_graph.m = mdp->m;
        g_graph.ns = mdp->ns;
        g_graph.gamma = mdp->gamma;
    }
    if (t0) hipEventRecord(t0, st);
    hipError_t e = hipGraphLaunch(g_graph.exec, st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to execute the sweep graph.", NOVA_ERROR_KERNEL_EXECUTION);
    if (t1) hipEventRecord(t1, st);
    e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to synchronize after the sweeps.",
                        NOVA_ERROR_DEVICE_SYNCHRONIZE);
    vi->currentHorizon += sweeps;
    return NOVA_SUCCESS;
}

int vi_sweeps_float(const nova::MDP *mdp, nova::MDPVIGPU *vi, unsigned sweeps, hipStream_t st,
                    const char *fn)
{
    return vi_run_sweeps(mdp, vi, sweeps, st, nullptr, nullptr, fn);
}

static bool mdp_device_ready(const nova::MDP *mdp)
{
    return mdp != nullptr && mdp->n != 0 && mdp->ns != 0 && mdp->m != 0 && mdp->d_S != nullptr &&
           mdp->d_T != nullptr && mdp->d_R != nullptr;
}

// mdp_vi_initialize_gpu's work: V, V' (from VInitial or zeros) and pi on the device, and, when
// layout, the sweep layout the model's size calls for.
static int vi_state_init(const nova::MDP *mdp, nova::MDPVIGPU *vi, const char *fn,
                         bool layout)
{
    if (mdp == nullptr || mdp->n == 0 || vi == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    sorted_drop(vi->d_V, fn);
    chunked_drop(vi->d_V, fn);
    device_free_if_owned(vi->d_V, fn, "the value function");
    device_free_if_owned(vi->d_VPrime, fn, "the value function (prime)");
    device_free_if_owned(vi->d_pi, fn, "the policy (pi)");
    vi->d_V = nullptr;
    vi->d_VPrime = nullptr;
    vi->d_pi = nullptr;
    vi->currentHorizon = 0;

    const size_t n = mdp->n;
    int rc = device_alloc((void **)&vi->d_V, n * sizeof(float), fn, "the value function");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc((void **)&vi->d_VPrime, n * sizeof(float), fn, "the value function (prime)");
    if (rc == NOVA_SUCCESS)
        rc = device_alloc((void **)&vi->d_pi, n * sizeof(unsigned), fn, "the policy (pi)");
    if (rc != NOVA_SUCCESS)
        return rc;

    std::vector<float> zeros;
    const float *V0 = vi->VInitial;
    if (V0 == nullptr) {
        zeros.assign(n, 0.0f);
        V0 = zeros.data();
    }
    hipError_t e = hipMemcpy(vi->d_V, V0, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(vi->d_VPrime, V0, n * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemset(vi->d_pi, 0, n * sizeof(unsigned));
    rc = hip_fail(e, fn, "Failed to copy memory from host to device for V/pi.",
                  NOVA_ERROR_MEMCPY_TO_DEVICE);
    if (layout && rc == NOVA_SUCCESS && mdp->d_S != nullptr && mdp->d_T != nullptr) {
        const int want = layout_wanted(mdp);
        if (want == 2)
            rc = sorted_build(mdp, vi->d_V, stream(), fn);
        else if (want == 3)
            rc = chunked_build(mdp, vi->d_V, stream(), fn);
    }
    return rc;
}

} // namespace nova_amd

namespace nova {

using namespace nova_amd;

int mdp_vi_initialize_gpu(const MDP *mdp, MDPVIGPU *vi)
{
    NOVA_AMD_API_LOCK();
    return vi_state_init(mdp, vi, "mdp_vi_initialize_gpu", true);
}

int mdp_vi_uninitialize_gpu(const MDP *mdp, MDPVIGPU *vi)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "mdp_vi_uninitialize_gpu";
    if (mdp == nullptr || vi == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    vi->currentHorizon = 0;
    int rc = NOVA_SUCCESS;
    sorted_drop(vi->d_V, fn);
    chunked_drop(vi->d_V, fn);
    if (device_free_if_owned(vi->d_V, fn, "the value function") != NOVA_SUCCESS)
        rc = NOVA_ERROR_DEVICE_FREE;
    if (device_free_if_owned(vi->d_VPrime, fn, "the value function (prime)") != NOVA_SUCCESS)
        rc = NOVA_ERROR_DEVICE_FREE;
    if (device_free_if_owned(vi->d_pi, fn, "the policy (pi)") != NOVA_SUCCESS)
        rc = NOVA_ERROR_DEVICE_FREE;
    vi->d_V = nullptr;
    vi->d_VPrime = nullptr;
    vi->d_pi = nullptr;
    return rc;
}

int mdp_vi_update_gpu(const MDP *mdp, MDPVIGPU *vi)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "mdp_vi_update_gpu";
    hipStream_t st = stream();
    const ViPlan plan = vi_plan(mdp, vi);
    const bool even = (vi->currentHorizon % 2) == 0;
    vi_launch(plan, mdp, even ? vi->d_V : vi->d_VPrime, even ? vi->d_VPrime : vi->d_V, vi->d_pi,
              true, st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to execute the 'Bellman update' kernel.",
                        NOVA_ERROR_KERNEL_EXECUTION);
    e = hipStreamSynchronize(st);
    if (e != hipSuccess)
        return hip_fail(e, fn, "Failed to synchronize the device after 'Bellman update' kernel.",
                        NOVA_ERROR_DEVICE_SYNCHRONIZE);
    vi->currentHorizon++;
    return NOVA_SUCCESS;
}

int mdp_vi_get_policy_gpu(const MDP *mdp, MDPVIGPU *vi, MDPValueFunction *policy)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "mdp_vi_get_policy_gpu";
    if (mdp == nullptr || vi == nullptr || policy == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (mdp_value_function_initialize(policy, mdp->n, mdp->m, 0) != NOVA_SUCCESS) {
        NOVA_ERR(fn, "Could not create the policy.");
        return NOVA_ERROR_POLICY_CREATION;
    }
    const float *dV = (vi->currentHorizon % 2 == 0) ? vi->d_V : vi->d_VPrime;
    hipError_t e = hipMemcpy(policy->V, dV, sizeof(float) * mdp->n, hipMemcpyDeviceToHost);
    if (e == hipSuccess)
        e = hipMemcpy(policy->pi, vi->d_pi, sizeof(unsigned) * mdp->n, hipMemcpyDeviceToHost);
    return hip_fail(e, fn, "Failed to copy memory from device to host for V/pi.",
                    NOVA_ERROR_MEMCPY_TO_HOST);
}

int mdp_vi_execute_gpu(const MDP *mdp, MDPVIGPU *vi, MDPValueFunction *policy)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "mdp_vi_execute_gpu";
    if (mdp == nullptr || mdp->n == 0 || mdp->ns == 0 || mdp->m == 0 || mdp->S == nullptr ||
            mdp->T == nullptr || mdp->R == nullptr || mdp->gamma < 0.0f || mdp->gamma > 1.0f ||
            mdp->horizon < 1 || vi == nullptr || policy == nullptr) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    if (vi->numThreads % 32 != 0) {
        NOVA_ERR(fn, "Invalid number of threads.");
        return NOVA_ERROR_INVALID_CUDA_PARAM;
    }
    if (!mdp_device_ready(mdp)) {
        NOVA_ERR(fn, "The MDP is not resident on the device (call mdp_initialize_gpu first).");
        return NOVA_ERROR_INVALID_DATA;
    }
    int rc = mdp_vi_initialize_gpu(mdp, vi);
    if (rc != NOVA_SUCCESS)
        return rc;
    rc = vi_run_sweeps(mdp, vi, mdp->horizon, stream(), nullptr, nullptr, fn);
    if (rc == NOVA_SUCCESS)
        rc = mdp_vi_get_policy_gpu(mdp, vi, policy);
    const int rc2 = mdp_vi_uninitialize_gpu(mdp, vi);
    return (rc != NOVA_SUCCESS) ? rc : rc2;
}

int mdp_vi_batch_execute_gpu(const MDP *const *mdps, MDPVIGPU *const *vis, unsigned int count,
                             MDPValueFunction *policies, float *elapsedMs)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "mdp_vi_batch_execute_gpu";
    if (mdps == nullptr || vis == nullptr || count == 0) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    for (unsigned b = 0; b < count; b++) {
        const MDP *mdp = mdps[b];
        if (!mdp_device_ready(mdp) || vis[b] == nullptr || mdp->horizon < 1 ||
                mdp->m > kBatchRows || mdp->gamma < 0.0f || mdp->gamma > 1.0f) {
            NOVA_ERR(fn, "Invalid instance (a device-resident MDP with m <= 1024, horizon >= 1, "
                         "gamma in [0, 1], and its solver struct).");
            return NOVA_ERROR_INVALID_DATA;
        }
    }
    int rc = NOVA_SUCCESS;
    std::vector<ViInst> host(count);
    std::vector<unsigned> off(count + 1, 0);
    unsigned maxH = 0;
    for (unsigned b = 0; b < count && rc == NOVA_SUCCESS; b++) {
        const MDP *mdp = mdps[b];
        rc = vi_state_init(mdp, vis[b], fn, false);
        ViInst &I = host[b];
        I.n = mdp->n;
        I.m = mdp->m;
        I.ns = mdp->ns;
        I.horizon = mdp->horizon;
        I.statesPerTile = kBatchRows / mdp->m;
        I.tiles = (mdp->n + I.statesPerTile - 1) / I.statesPerTile;
        I.gamma = mdp->gamma;
        I.S = mdp->d_S;
        I.T = mdp->d_T;
        I.R = mdp->d_R;
        I.V = vis[b]->d_V;
        I.Vp = vis[b]->d_VPrime;
        I.pi = vis[b]->d_pi;
        off[b + 1] = off[b] + I.tiles;
        maxH = std::max(maxH, I.horizon);
        if ((size_t)off[b] + I.tiles >= ((size_t)1 << 31))
            rc = NOVA_ERROR_INVALID_DATA;
    }
    void *mem = nullptr;
    const size_t bI = up256(sizeof(ViInst) * count), bO = up256(sizeof(unsigned) * (count + 1));
    if (rc == NOVA_SUCCESS)
        rc = device_alloc(&mem, bI + bO, fn, "the batch descriptors");
    hipStream_t st = stream();
    hipGraphExec_t exec = nullptr;
    if (rc == NOVA_SUCCESS) {
        ViInst *dInst = static_cast<ViInst *>(mem);
        unsigned *dOff = reinterpret_cast<unsigned *>(static_cast<char *>(mem) + bI);
        hipError_t e = hipMemcpyAsync(dInst, host.data(), sizeof(ViInst) * count,
                                      hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(dOff, off.data(), sizeof(unsigned) * (count + 1),
                               hipMemcpyHostToDevice, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        hipStream_t cs = capture_stream();
        hipGraph_t graph = nullptr;
        if (e == hipSuccess && cs == nullptr)
            e = hipErrorInvalidValue;
        if (e == hipSuccess)
            e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
        if (e == hipSuccess) {
            for (unsigned k = 0; k < maxH; k++)
                hipLaunchKernelGGL(vi_batch_sweep, dim3(off[count]), dim3(kBatchThreads), 0, cs,
                                   dInst, dOff, count, k);
            e = hipStreamEndCapture(cs, &graph);
        }
        if (e == hipSuccess) {
            e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            hipGraphDestroy(graph);
        }
        hipEvent_t t0 = nullptr, t1 = nullptr;
        if (e == hipSuccess && elapsedMs) {
            hipEventCreate(&t0);
            hipEventCreate(&t1);
            hipEventRecord(t0, st);
        }
        if (e == hipSuccess)
            e = hipGraphLaunch(exec, st);
        if (e == hipSuccess && elapsedMs)
            hipEventRecord(t1, st);
        if (e == hipSuccess)
            e = hipStreamSynchronize(st);
        if (elapsedMs) {
            float ms = 0.0f;
            if (e == hipSuccess && t0 && t1)
                hipEventElapsedTime(&ms, t0, t1);
            *elapsedMs = ms;
        }
        if (t0) hipEventDestroy(t0);
        if (t1) hipEventDestroy(t1);
        if (exec)
            hipGraphExecDestroy(exec);
        rc = hip_fail(e, fn, "Failed to run the batch graph.", NOVA_ERROR_KERNEL_EXECUTION);
    }
    if (mem)
        device_free(mem, fn, "the batch descriptors");
    for (unsigned b = 0; b < count && rc == NOVA_SUCCESS; b++) {
        vis[b]->currentHorizon = mdps[b]->horizon;
        if (policies)
            rc = mdp_vi_get_policy_gpu(mdps[b], vis[b], &policies[b]);
    }
    return rc;
}

int mdp_vi_sweep_kernel_gpu(const MDP *mdp, const MDPVIGPU *vi, unsigned int *kernel)
{
    NOVA_AMD_API_LOCK();
    if (mdp == nullptr || vi == nullptr || kernel == nullptr) {
        NOVA_ERR("mdp_vi_sweep_kernel_gpu", "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    const ViPlan p = vi_plan(mdp, vi);
    *kernel = p.chunked ? 3u : (p.sorted ? 2u : (p.tiled ? 1u : 0u));
    return NOVA_SUCCESS;
}

int mdp_vi_run_gpu(const MDP *mdp, MDPVIGPU *vi, const float *d_V0, unsigned int sweeps,
                   float *elapsedMs)
{
    NOVA_AMD_API_LOCK();
    const char *fn = "mdp_vi_run_gpu";
    if (!mdp_device_ready(mdp) || vi == nullptr || vi->d_V == nullptr ||
            vi->d_VPrime == nullptr || vi->d_pi == nullptr || sweeps == 0) {
        NOVA_ERR(fn, "Invalid arguments.");
        return NOVA_ERROR_INVALID_DATA;
    }
    hipStream_t st = stream();
    if (d_V0 != nullptr) {
        const size_t bytes = sizeof(float) * mdp->n;
        hipError_t e = hipMemcpyAsync(vi->d_V, d_V0, bytes, hipMemcpyDeviceToDevice, st);
        if (e == hipSuccess)
            e = hipMemcpyAsync(vi->d_VPrime, d_V0, bytes, hipMemcpyDeviceToDevice, st);
        if (e != hipSuccess)
            return hip_fail(e, fn, "Failed to reset V from d_V0.", NOVA_ERROR_MEMCPY_TO_DEVICE);
        vi->currentHorizon = 0;
    }
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (elapsedMs) {
        hipEventCreate(&t0);
        hipEventCreate(&t1);
    }
    int rc = vi_run_sweeps(mdp, vi, sweeps, st, t0, t1, fn);
    if (elapsedMs) {
        float ms = 0.0f;
        if (rc == NOVA_SUCCESS)
            hipEventElapsedTime(&ms, t0, t1);
        *elapsedMs = ms;
        hipEventDestroy(t0);
        hipEventDestroy(t1);
    }
    return rc;
}

} // namespace nova
