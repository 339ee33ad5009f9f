// shk_sliced.hpp -- the sliced SpMV layout and its kernel (RowSparse).
//
// Layout.  Rows are taken in panels of kSlRows = 8192 consecutive rows, one
// 512-thread work-group (8 wavefronts) per panel.  Inside a panel the rows
// are ordered by length (longest first) and cut into groups of 64 -- one row
// per lane -- and every wavefront owns 16 groups, so every lane owns 16 rows
// for the whole product and keeps their running sums in registers.  The
// groups are dealt to the wavefronts so that their per-chunk work evens out
// (the work-group waits for its slowest wavefront at every chunk).
//
// A panel's entries are cut into chunks by column windows of at most
// sl_win<T>() columns (64 KiB of x), staged in LDS once per chunk.  Within a
// chunk each row owns one segment (its entries in the window: a contiguous
// run, the row's columns being sorted).  Per chunk a wavefront walks its 16
// groups in order; group i takes K_i steps (its longest segment), step k
// holding the k-th segment entry of every lane whose segment is longer than
// k.  Those entries are stored at consecutive positions in lane order, so
// the lanes active at a step read consecutive stream positions (pos +
// mbcnt(active)): coalesced loads; and every lane adds its own row's
// products in stored order -- Seldon's `temp += data[j] * X(ind[j])`
// (Functions_MatVect.cxx:424-430) -- bitwise like the CPU.  No product is
// staged in LDS: the only LDS traffic is the x window and one x read per
// entry.  An entry is stored as its window-local column (16 bits) and its
// value.
//
// A wavefront's steps, chunk after chunk, form one stream cut into blocks
// of kSlRing = 16 steps (a chunk's steps are padded to whole blocks).  Per
// block every lane has 16 bits: bit r set when it is active at step r.  Per
// (chunk, wavefront) a descriptor holds the block count, the 16 groups' step
// counts, and the window (c0, width); per (panel, wavefront) the stream's
// first position and first block.
//
// Kernel.  The producer runs one block (16 steps) ahead of the consumer: it
// turns a block's bits into the steps' active masks and issues their loads
// into a ring of 16 register slots; the consumer reads x from the window,
// multiplies and adds in order.  The producer needs nothing but the bits, so
// its loads are the same every block (no branch: the wait counters stay
// exact).  The consumer switches accumulators where a group ends (a uniform
// index into the register array).  Entries are read with buffer loads (the
// step's first position in a scalar offset, the lane's rank among the active
// lanes in the vector offset), and the window is staged with loads that
// write LDS directly, so no registers hold it.
#pragma once

namespace shk {

constexpr int kSlThreads = 512;
constexpr int kSlWaves = kSlThreads / kWave;  // 8
constexpr int kSlGroups = 16;                 // row groups per wavefront
constexpr int kSlRows = kSlWaves * kSlGroups * kWave;  // 8192 rows per panel
constexpr int kSlRing = 16;                   // steps per block: 16 activity bits per lane
constexpr unsigned short kSlNoRow = 0xffffu;
template <typename T>
constexpr int sl_win() { return sizeof(T) == 8 ? 8192 : 16384; }  // 64 KiB of x in LDS
static_assert(sl_win<double>() >= kSlRows && sl_win<float>() >= kSlRows, "a panel's sums fit the window");

// descriptor of one (chunk, wavefront): {blocks, group step counts (bytes
// 0-15), window start c0, window width}
struct alignas(32) SlDesc {
    int nblk;
    unsigned k[4];
    int c0, width, unused;
};

struct SlicedView {
    int npanel;
    int ncol;                        // x's length (bound of the window loads)
    const int *pchunk;               // [npanel + 1] first chunk of each panel
    const SlDesc *desc;              // [nchunk * kSlWaves]
    const int2 *wstart;              // [npanel * kSlWaves] {first stream position, first block}
    const unsigned short *bits;      // [blocks * kWave] activity bits of each block and lane
    int bits_bytes;                  // size of bits (buffer bound)
    const unsigned short *rowmap;    // [npanel * kSlWaves * kSlGroups * kWave] panel-local row (kSlNoRow: none)
    const unsigned short *col;       // entry stream: LDS byte offset of the entry's x in the window
    int stream_len;                  // stream length with its padding (buffer bound, entries)
};

// byte i of the 16 step counts (a select chain: an indexed read of the
// array would put the descriptor in scratch memory)
__host__ __device__ __forceinline__ int sl_byte(const unsigned (&k)[4], int i)
{
    const int q = i >> 2;
    const unsigned v = q == 0 ? k[0] : q == 1 ? k[1] : q == 2 ? k[2] : k[3];
    return (int)((v >> ((i & 3) * 8)) & 0xffu);
}

// raw buffer resource over n bytes (gfx9 word 3)
__device__ __forceinline__ __amdgpu_buffer_rsrc_t sl_rsrc(const void *p, int nbytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, nbytes, 0x00020000);
}

// Cache policy of the entry and bit streams: non-temporal (each byte is read
// once) for FP64, where it measured 6% faster; the default for FP32, where
// it made no difference.
template <typename T>
constexpr int sl_aux() { return sizeof(T) == 8 ? 2 : 0; }

template <typename T>
__device__ __forceinline__ T sl_load(__amdgpu_buffer_rsrc_t r, int voff, int soff);
template <>
__device__ __forceinline__ double sl_load<double>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, sl_aux<double>()));
}
template <>
__device__ __forceinline__ float sl_load<float>(__amdgpu_buffer_rsrc_t r, int voff, int soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, sl_aux<float>()));
}

template <typename T>
__global__ __launch_bounds__(kSlThreads, 4) void k_sliced(SpmvArgs<T> a, SlicedView s)
{
    if (a.skip && *a.skip) return;
    __shared__ T xs[sl_win<T>()];
    const int p = xcd_logical_block(a.xcd_q);
    const int t = threadIdx.x, lane = t & (kWave - 1);
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    T c[2] = {(T)0, (T)0};
    if (p < s.npanel) {
        const int cb = s.pchunk[p], ce = s.pchunk[p + 1];
        T acc[kSlGroups];
#pragma unroll
        for (int i = 0; i < kSlGroups; i++) acc[i] = (T)0;
        const __amdgpu_buffer_rsrc_t rcol = sl_rsrc(s.col, s.stream_len * 2);
        const __amdgpu_buffer_rsrc_t rval = sl_rsrc(a.val, s.stream_len * (int)sizeof(T));
        const __amdgpu_buffer_rsrc_t rbits = sl_rsrc(s.bits, s.bits_bytes);
        const __amdgpu_buffer_rsrc_t rx = sl_rsrc(a.x, s.ncol * (int)sizeof(T));
        const int2 ws = s.wstart[(size_t)p * kSlWaves + w];
        int ppos = ws.x;
        int bofs = (ws.y * kWave + lane) * 2;  // byte offset of this lane's bits of the next block to load
        // ring slots: value, LDS byte offset of x, lane active, step's active mask
        T rv[kSlRing];
        unsigned short rw[kSlRing];  // 16-bit: widened where used, not at the loop's end
        bool ra[kSlRing];
        uint64_t rm[kSlRing];
        // bits of the block being produced, of the one after it, and of the
        // block after that (loaded at the start of a block, before its entry
        // loads, so the rotation at the block's end never waits for them)
        unsigned bcur, bnext, bload = 0;
        bcur = __builtin_amdgcn_raw_buffer_load_b16(rbits, bofs, 0, 0);
        bnext = __builtin_amdgcn_raw_buffer_load_b16(rbits, bofs + kWave * 2, 0, 0);
        bofs += 2 * kWave * 2;
        // step r of the next block into slot r (every lane loads; an inactive
        // lane reads the entry of the next active lane above it, or past the
        // step: in the stream's padding at worst, never used)
        auto produce = [&](int r) {
            const bool act = (bcur >> r) & 1u;
            const uint64_t m = __builtin_amdgcn_ballot_w64(act);
            const int below =
                (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
            rw[r] = __builtin_amdgcn_raw_buffer_load_b16(rcol, below * 2, ppos * 2, sl_aux<T>());
            rv[r] = sl_load<T>(rval, below * (int)sizeof(T), ppos * (int)sizeof(T));
            ra[r] = act;
            rm[r] = m;
            ppos += __builtin_popcountll(m);
        };
        auto load_bits = [&]() {
            bload = __builtin_amdgcn_raw_buffer_load_b16(rbits, bofs, 0, sl_aux<T>());
            bofs += kWave * 2;
        };
        auto next_bits = [&]() {
            bcur = bnext;
            bnext = bload;
        };
        load_bits();
#pragma unroll
        for (int r = 0; r < kSlRing; r++) produce(r);
        next_bits();
        constexpr int VPL = 16 / (int)sizeof(T);  // x values per lane per 16-byte window load
        for (int cc = cb; cc < ce; cc++) {
            const SlDesc cd = s.desc[(size_t)cc * kSlWaves + w];
            // stage the window x[c0, c0 + sl_win) straight into LDS: wave w
            // fills its eighth, 1 KiB per load, as far as the window's width
            // (reads past x's end return 0; entries past the width are never
            // read)
            __syncthreads();  // the previous window's reads are done
#pragma unroll
            for (int j = 0; j < sl_win<T>() / (kSlWaves * kWave * VPL); j++) {
                const int e0 = (w * (sl_win<T>() / kSlWaves)) + j * kWave * VPL;  // first element of this load
                if (e0 >= cd.width) break;  // wave-uniform: the rest of this wave's share is past the window
                __builtin_amdgcn_raw_ptr_buffer_load_lds(
                    rx, (__attribute__((address_space(3))) void *)(xs + e0), 16,
                    (cd.c0 + e0 + lane * VPL) * (int)sizeof(T), 0, 0, 0);
            }
            __builtin_amdgcn_s_waitcnt(0x0f70);  // vmcnt(0): the window has landed in LDS
            __syncthreads();
            if (cd.nblk == 0) continue;
            // the chunk's first group with steps; its accumulator in `cur`
            int gi = 0;
            while (sl_byte(cd.k, gi) == 0) gi++;
            int gend = sl_byte(cd.k, gi);
            T cur = acc[gi];
            int step = 0;
            for (int b = 0; b < cd.nblk; b++) {
                load_bits();
#pragma unroll
                for (int r = 0; r < kSlRing; r++) {
                    if (step == gend) {  // group gi ends: the next one with steps
                        acc[gi] = cur;
                        do {
                            gi++;
                        } while (gi < kSlGroups && sl_byte(cd.k, gi) == 0);
                        if (gi < kSlGroups) {
                            cur = acc[gi];
                            gend += sl_byte(cd.k, gi);
                        } else {
                            gi = kSlGroups - 1;  // only padding steps are left (no lane active)
                            cur = acc[gi];
                            gend = INT32_MAX;
                        }
                    }
                    if (rm[r]) {
                        const T xv = *(const T *)((const char *)xs + (unsigned)rw[r]);
                        const T pr = rv[r] * xv;
                        const T sum = cur + pr;
                        cur = ra[r] ? sum : cur;
                    }
                    produce(r);
                    step++;
                }
                next_bits();
            }
            acc[gi] = cur;
        }
        // y = beta y + alpha temp (Functions_MatVect.cxx:414, :430) in row
        // order: the sums go to LDS at their panel rows (the window's space
        // holds a panel's 8192 values), then y is written coalesced instead
        // of at the lanes' length-sorted rows, a partial cache line per lane
        const unsigned short *rmap = s.rowmap + ((size_t)p * kSlWaves + w) * kSlGroups * kWave + lane;
        __syncthreads();  // the last window's reads are done
#pragma unroll
        for (int i = 0; i < kSlGroups; i++) {
            const unsigned short lr = rmap[i * kWave];
            if (lr != kSlNoRow) xs[lr] = acc[i];
        }
        __syncthreads();
        const int nr = min(kSlRows, a.m - p * kSlRows);  // every panel row has a slot
        for (int lr = t; lr < nr; lr += kSlThreads) {
            const int row = p * kSlRows + lr;
            T y = panel_y0<T, kCsr>(a, row);
            y += a.alpha * xs[lr];
            spmv_epilogue<T, kCsr>(a, row, y, c);
        }
    }
    if (a.dot0 != kDotNone || a.dot1 != kDotNone) block_partial<T, 2>(c, a.partials, gridDim.x, blockIdx.x);
}

}  // namespace shk
