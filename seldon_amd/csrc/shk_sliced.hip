// shk_sliced.hip -- host side of the sliced SpMV layout (RowSparse): the
// planner, the upload, value updates.  The layout and its kernel are
// described in shk_sliced.hpp.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>

#include "shk_kernels.hpp"

namespace shk {

namespace {

// entries past the stream's end (read, never used)
constexpr int kSlPad = 256;

struct SlPanel {
    int nch = 0;
    std::vector<int2> win;              // per chunk {c0, c1}
    std::vector<SlDesc> desc;           // per chunk and wavefront
    std::vector<int2> wstart;           // per wavefront {stream start, first block}, relative to the panel
    std::vector<unsigned short> bits;   // per block and lane: 16 activity bits
    std::vector<unsigned short> rowmap; // [kSlWaves][kSlGroups][kWave]
    std::vector<unsigned short> col;
    std::vector<int> src;
    double x_cols = 0;
    bool ok = true;
};

// One panel: rows [r0, r1) of the caller's CSR (sorted columns checked by
// the caller).
void plan_panel(const int *ptr, const int *ind, int n, int r0, int r1, int win, int vsize, SlPanel &P)
{
    const int nr = r1 - r0;
    std::vector<int> order(nr);
    for (int k = 0; k < nr; k++) order[k] = k;
    // slots by length, longest first: a group of 64 slots holds rows of
    // similar length, so its steps keep most lanes busy
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return ptr[r0 + a + 1] - ptr[r0 + a] > ptr[r0 + b + 1] - ptr[r0 + b];
    });
    std::vector<int> cur(nr), end(nr);
    for (int s = 0; s < nr; s++) {
        cur[s] = ptr[r0 + order[s]];
        end[s] = ptr[r0 + order[s] + 1];
    }
    // column windows: from the smallest column left, at most `win` columns,
    // cut early so no row has more than 255 entries in one window
    std::vector<std::vector<uint8_t>> seg;   // per chunk, per slot
    std::vector<std::vector<int>> segst;     // per chunk, per slot: first entry
    for (;;) {
        int c0 = INT32_MAX;
        for (int s = 0; s < nr; s++)
            if (cur[s] < end[s]) c0 = std::min(c0, ind[cur[s]]);
        if (c0 == INT32_MAX) break;
        int c1 = (int)std::min<int64_t>((int64_t)c0 + win, n);
        for (int s = 0; s < nr; s++)
            if (cur[s] + 255 < end[s]) c1 = std::min(c1, ind[cur[s] + 255]);
        if (c1 <= c0) {  // 256 entries of one row in a single column
            P.ok = false;
            return;
        }
        std::vector<uint8_t> len(kSlRows, 0);
        std::vector<int> st(kSlRows, 0);
        for (int s = 0; s < nr; s++) {
            int j = cur[s];
            st[s] = j;
            while (j < end[s] && ind[j] < c1) j++;
            len[s] = (uint8_t)(j - cur[s]);
            cur[s] = j;
        }
        P.win.push_back(make_int2(c0, c1));
        P.x_cols += c1 - c0;
        seg.push_back(std::move(len));
        segst.push_back(std::move(st));
    }
    const int nch = (int)P.win.size();
    P.nch = nch;
    constexpr int NG = kSlRows / kWave;  // 128 groups
    // step counts per chunk and group (the group's longest segment)
    std::vector<int> K((size_t)nch * NG, 0);
    std::vector<int64_t> tot(NG, 0);
    for (int c = 0; c < nch; c++)
        for (int g = 0; g < NG; g++) {
            int k = 0;
            for (int l = 0; l < kWave; l++) k = std::max(k, (int)seg[c][g * kWave + l]);
            K[(size_t)c * NG + g] = k;
            tot[g] += k;
        }
    // groups to wavefronts, 16 each: longest total first, each to the
    // wavefront whose largest per-chunk load it raises least (the work-group
    // waits for its slowest wavefront at every chunk)
    std::vector<int> gorder(NG);
    for (int g = 0; g < NG; g++) gorder[g] = g;
    std::stable_sort(gorder.begin(), gorder.end(), [&](int a, int b) { return tot[a] > tot[b]; });
    std::vector<int64_t> load((size_t)kSlWaves * nch, 0), wtot(kSlWaves, 0);
    std::vector<int> cnt(kSlWaves, 0), grp((size_t)kSlWaves * kSlGroups, 0);
    for (int g : gorder) {
        int best = -1;
        int64_t bmax = 0, btot = 0;
        for (int w = 0; w < kSlWaves; w++) {
            if (cnt[w] == kSlGroups) continue;
            int64_t mx = 0;
            for (int c = 0; c < nch; c++) mx = std::max(mx, load[(size_t)w * nch + c] + K[(size_t)c * NG + g]);
            if (best < 0 || mx < bmax || (mx == bmax && wtot[w] < btot)) {
                best = w;
                bmax = mx;
                btot = wtot[w];
            }
        }
        for (int c = 0; c < nch; c++) load[(size_t)best * nch + c] += K[(size_t)c * NG + g];
        wtot[best] += tot[g];
        grp[(size_t)best * kSlGroups + cnt[best]++] = g;
    }
    for (int w = 0; w < kSlWaves; w++)
        std::sort(grp.begin() + (size_t)w * kSlGroups, grp.begin() + (size_t)(w + 1) * kSlGroups);
    P.rowmap.assign((size_t)kSlWaves * kSlGroups * kWave, kSlNoRow);
    for (int w = 0; w < kSlWaves; w++)
        for (int i = 0; i < kSlGroups; i++)
            for (int l = 0; l < kWave; l++) {
                const int s = grp[(size_t)w * kSlGroups + i] * kWave + l;
                if (s < nr) P.rowmap[((size_t)w * kSlGroups + i) * kWave + l] = (unsigned short)order[s];
            }
    // the streams, wavefront after wavefront, chunk after chunk: group after
    // group, step k = the k-th segment entry of each lane with a longer
    // segment; each chunk's steps padded to whole blocks
    P.desc.resize((size_t)nch * kSlWaves);
    P.wstart.resize(kSlWaves);
    int nblk = 0;
    for (int w = 0; w < kSlWaves; w++) {
        P.wstart[w] = make_int2((int)P.col.size(), nblk);
        for (int c = 0; c < nch; c++) {
            const int c0 = P.win[c].x;
            SlDesc &d = P.desc[(size_t)c * kSlWaves + w];
            std::memset(&d, 0, sizeof(d));
            d.c0 = c0;
            d.width = P.win[c].y - c0;
            int steps = 0;
            for (int i = 0; i < kSlGroups; i++) steps += K[(size_t)c * NG + grp[(size_t)w * kSlGroups + i]];
            d.nblk = (steps + kSlRing - 1) / kSlRing;
            const size_t b0 = P.bits.size();
            P.bits.resize(b0 + (size_t)d.nblk * kWave, 0);
            int step = 0;
            for (int i = 0; i < kSlGroups; i++) {
                const int g = grp[(size_t)w * kSlGroups + i];
                const int k = K[(size_t)c * NG + g];
                d.k[i >> 2] |= (unsigned)k << ((i & 3) * 8);
                for (int q = 0; q < k; q++, step++)
                    for (int l = 0; l < kWave; l++) {
                        const int s = g * kWave + l;
                        if (seg[c][s] > q) {
                            const int j = segst[c][s] + q;
                            // the LDS byte offset of x[ind[j]] in the window
                            P.col.push_back((unsigned short)((ind[j] - c0) * vsize));
                            P.src.push_back(j);
                            P.bits[b0 + (size_t)(step / kSlRing) * kWave + l] |=
                                (unsigned short)(1u << (step % kSlRing));
                        }
                    }
            }
            nblk += d.nblk;
        }
    }
}

struct SlPlan {
    int npanel = 0;
    std::vector<int> pchunk;
    std::vector<SlDesc> desc;
    std::vector<int2> wstart;
    std::vector<unsigned short> bits;
    std::vector<unsigned short> rowmap;
    std::vector<unsigned short> col;
    std::vector<int> src;
};

// false when the layout does not apply: unsorted columns in a row, or x
// windows costing more traffic than the matrix itself
bool plan_sliced(int m, int n, const int *ptr, const int *ind, int win, size_t vsize, SlPlan &Q)
{
    if (m <= 0) return false;
    for (int r = 0; r < m; r++)
        for (int j = ptr[r] + 1; j < ptr[r + 1]; j++)
            if (ind[j] < ind[j - 1]) return false;
    const int64_t nnz = ptr[m];
    Q.npanel = (m + kSlRows - 1) / kSlRows;
    std::vector<SlPanel> P(Q.npanel);
    const int nth = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (int k = 0; k < nth; k++)
        th.emplace_back([&, k]() {
            for (int p = k; p < Q.npanel; p += nth)
                plan_panel(ptr, ind, n, p * kSlRows, std::min(m, (p + 1) * kSlRows), win, (int)vsize, P[p]);
        });
    for (auto &t : th) t.join();
    double xcols = 0;
    size_t ns = 0, nch = 0, nbits = 0;
    for (const auto &pp : P) {
        if (!pp.ok) return false;
        xcols += pp.x_cols;
        ns += pp.col.size();
        nch += pp.win.size();
        nbits += pp.bits.size();
    }
    if (xcols * (double)vsize > 1.5 * (double)nnz * (double)(4 + vsize)) return false;
    // buffer loads address the stream, the bits and x with 32-bit byte offsets
    if ((int64_t)ns != nnz || (ns + kSlPad) * vsize >= (size_t)INT32_MAX ||
        (nbits + 4 * kWave) * 2 >= (size_t)INT32_MAX || (size_t)n * vsize >= (size_t)INT32_MAX)
        return false;
    Q.pchunk.assign(Q.npanel + 1, 0);
    Q.desc.reserve(nch * kSlWaves);
    Q.wstart.reserve((size_t)Q.npanel * kSlWaves);
    Q.bits.reserve(nbits + 4 * kWave);
    Q.rowmap.reserve((size_t)Q.npanel * kSlWaves * kSlGroups * kWave);
    Q.col.reserve(ns + kSlPad);
    Q.src.reserve(ns);
    for (int p = 0; p < Q.npanel; p++) {
        SlPanel &pp = P[p];
        const int base = (int)Q.col.size(), bbase = (int)(Q.bits.size() / kWave);
        for (int2 v : pp.wstart) Q.wstart.push_back(make_int2(v.x + base, v.y + bbase));
        Q.desc.insert(Q.desc.end(), pp.desc.begin(), pp.desc.end());
        Q.bits.insert(Q.bits.end(), pp.bits.begin(), pp.bits.end());
        Q.rowmap.insert(Q.rowmap.end(), pp.rowmap.begin(), pp.rowmap.end());
        Q.col.insert(Q.col.end(), pp.col.begin(), pp.col.end());
        Q.src.insert(Q.src.end(), pp.src.begin(), pp.src.end());
        Q.pchunk[p + 1] = Q.pchunk[p] + pp.nch;
        SlPanel().bits.swap(pp.bits);
        std::vector<unsigned short>().swap(pp.col);
        std::vector<int>().swap(pp.src);
    }
    // the producer reads up to three blocks past a wavefront's last one and
    // re-reads the stream at its end: zero blocks and padding entries
    Q.bits.resize(Q.bits.size() + 4 * kWave, 0);
    Q.col.resize(ns + kSlPad, 0);
    return true;
}

template <typename T>
__global__ void k_sl_gather(int64_t ns, const int *src, const T *from, T *to)
{
    for (int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x; q < ns; q += (int64_t)gridDim.x * kBlock)
        to[q] = from[src[q]];
}

template <typename V>
int put(V **dst, const std::vector<V> &src, hipStream_t s)
{
    SHK_HIP(hipMalloc((void **)dst, sizeof(V) * std::max<size_t>(src.size(), 1)));
    if (!src.empty()) SHK_HIP(hipMemcpyAsync(*dst, src.data(), sizeof(V) * src.size(), hipMemcpyHostToDevice, s));
    return SHK_OK;
}

}  // namespace

int sliced_build(shk_csr A, const int *ptr, const int *ind, bool *applies)
{
    *applies = false;
    SlPlan Q;
    const size_t vs = dsize(A->dtype);
    const int win = A->dtype == SHK_F64 ? sl_win<double>() : sl_win<float>();
    if (A->storage != SHK_ROWSPARSE || !plan_sliced(A->m, A->n, ptr, ind, win, vs, Q)) return SHK_OK;
    hipStream_t s = A->ctx->stream;
    A->sl_npanel = Q.npanel;
    A->sl_nchunk = Q.pchunk[Q.npanel];
    A->sl_q = (Q.npanel + kNumXcd - 1) / kNumXcd;
    A->sl_launch = A->sl_q * kNumXcd;
    A->sl_ns = (int64_t)Q.src.size();
    SHK_TRY(put(&A->sl_pchunk, Q.pchunk, s));
    SHK_TRY(put((SlDesc **)&A->sl_desc, Q.desc, s));
    SHK_TRY(put((int2 **)&A->sl_wstart, Q.wstart, s));
    SHK_TRY(put((unsigned short **)&A->sl_bits, Q.bits, s));
    A->sl_bits_bytes = (int)(Q.bits.size() * sizeof(unsigned short));
    A->sl_stream_len = (int)Q.col.size();
    SHK_TRY(put(&A->sl_rowmap, Q.rowmap, s));
    SHK_TRY(put(&A->sl_col, Q.col, s));
    SHK_TRY(put(&A->sl_src, Q.src, s));
    const size_t nv = (size_t)A->sl_ns + kSlPad;
    SHK_HIP(hipMalloc(&A->sl_val, vs * nv));
    SHK_HIP(hipMemsetAsync(A->sl_val, 0, vs * nv, s));
    SHK_HIP(hipStreamSynchronize(s));  // host vectors go out of scope
    *applies = true;
    return SHK_OK;
}

// the stream's values from the row layout's (the caller's order)
int sliced_set_values(shk_csr A)
{
    if (!A->sl_val || A->sl_ns == 0) return SHK_OK;
    hipStream_t s = A->ctx->stream;
    if (A->dtype == SHK_F64)
        hipLaunchKernelGGL(k_sl_gather<double>, dim3(red_blocks(A->sl_ns)), dim3(kBlock), 0, s, A->sl_ns,
                           (const int *)A->sl_src, (const double *)A->val, (double *)A->sl_val);
    else
        hipLaunchKernelGGL(k_sl_gather<float>, dim3(red_blocks(A->sl_ns)), dim3(kBlock), 0, s, A->sl_ns,
                           (const int *)A->sl_src, (const float *)A->val, (float *)A->sl_val);
    SHK_HIP(hipGetLastError());
    return SHK_OK;
}

void sliced_free(shk_csr A)
{
    hipFree(A->sl_pchunk);
    hipFree(A->sl_desc);
    hipFree(A->sl_wstart);
    hipFree(A->sl_bits);
    hipFree(A->sl_rowmap);
    hipFree(A->sl_col);
    hipFree(A->sl_val);
    hipFree(A->sl_src);
    A->sl_pchunk = nullptr;
    A->sl_desc = A->sl_wstart = A->sl_bits = A->sl_val = nullptr;
    A->sl_rowmap = A->sl_col = nullptr;
    A->sl_src = nullptr;
}

}  // namespace shk
