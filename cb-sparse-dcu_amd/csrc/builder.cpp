// CB builder: row-major COO -> blocks, column gathering, per-block format,
// row-owned segments or LPT balance, flat CSR and row patterns.
//
// Output is bit-identical to the reference builder (coo2block_gather,
// cb-spmv-dcu/src/coo2block.h:628-670, and coo2block_comp, :605-626) for the
// same options.  The reference walks the whole COO stream serially with hash
// maps and one std::vector per block; this builder relies on the row-major
// input contract (rows non-decreasing, which makes every block row a
// contiguous slice of the stream) to work block row by block row in parallel:
//   pass 1  per block row: raw block counts in original column space
//           (first-appearance numbering is per block row because block ids of
//           a row-major stream are ordered by block row) -> gather decision
//   pass 2  per block row: gathered column list, block numbering in the
//           (gathered) column space, entries grouped per block in input order
//   pass 3  per block: format choice and payload encoding
// Tie orders that the reference leaves to libstdc++'s unstable std::sort (the
// COO entry order of duplicates, coo2block.h:219-227; the LPT order of blocks
// with equal nnz, :409) are reproduced by calling std::sort with an equivalent
// comparator on the same input sequence.
#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <queue>
#include <unordered_map>

#include <omp.h>

#include "cb_plan.hpp"

namespace cbs {
namespace {

int trunc_orig(int bs, int v) { return bs <= 16 ? (v & 0xff) : (v & 0xffff); }

// Per-thread scratch for first-appearance numbering of block columns.
struct Numbering {
    std::vector<int> stamp, id;
    void ensure(int gb_n) {
        if ((int)stamp.size() < gb_n) { stamp.assign(gb_n, -1); id.assign(gb_n, 0); }
    }
};

// 64-bit mix for row-pattern hashing (splitmix64 finaliser).
inline uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

double group_stddev(const std::vector<long long>& loads) {
    double sum = 0.0;
    for (long long l : loads) sum += (double)l;
    const double mean = sum / (double)loads.size();
    double var = 0.0;
    for (long long l : loads) { const double d = (double)l - mean; var += d * d; }
    return std::sqrt(var / (double)loads.size());
}

// ---- LPT balance (coo2block.h:382-498) -------------------------------------
struct BlockIdx { int blk_nnz, origin_idx, end_idx; };

void balance(Plan& p) {
    const int W = p.o.waves;
    std::vector<BlockIdx> a(p.nnzb);
    for (int i = 0; i < p.nnzb; ++i) a[i] = {p.nnzpb[i], i, 0};
    std::vector<long long> before((p.nnzb + W - 1) / W, 0);
    for (int i = 0; i < p.nnzb; ++i) before[i / W] += p.nnzpb[i];
    const double sdb = group_stddev(before);

    std::sort(a.begin(), a.end(), [](BlockIdx x, BlockIdx y) { return x.blk_nnz > y.blk_nnz; });
    int groups = p.nnzb / W;
    if (p.nnzb % W != 0) {
        groups += 1;
        const int pad = W - p.nnzb % W;
        for (int q = 0; q < pad; ++q) a.push_back({0, INT_MAX, -1});
        p.nnzb += pad;
    }
    // groups kept in a binary min-heap on (load, group index)
    struct Group { long long total; int index, count; };
    auto after_cmp = [](const Group& x, const Group& y) {
        return x.total != y.total ? x.total > y.total : x.index > y.index;
    };
    std::priority_queue<Group, std::vector<Group>, decltype(after_cmp)> heap(after_cmp);
    for (int g = 0; g < groups; ++g) heap.push({0, g, 0});
    for (int i = 0; i < p.nnzb; ++i) {
        Group g = heap.top();
        heap.pop();
        a[i].end_idx = g.index * W + g.count;
        g.total += a[i].blk_nnz;
        if (++g.count < W) heap.push(g);
    }
    std::sort(a.begin(), a.end(), [](BlockIdx x, BlockIdx y) { return x.end_idx < y.end_idx; });

    std::vector<int> nr(p.nnzb), nc(p.nnzb), nn(p.nnzb);
    std::vector<uint8_t> nt(p.nnzb);
    std::vector<long long> after(groups, 0);
    for (int i = 0; i < p.nnzb; ++i) {
        const int o = a[i].origin_idx;
        const bool real = o != INT_MAX;
        nr[i] = real ? p.brow[o] : -1;
        nc[i] = real ? p.bcol[o] : -1;
        nn[i] = real ? p.nnzpb[o] : 0;
        nt[i] = real ? p.type[o] : kCOO;
        after[i / W] += a[i].blk_nnz;
    }
    p.balanced = true;
    p.sd_before = sdb;
    p.sd_after = group_stddev(after);
    p.brow.swap(nr); p.bcol.swap(nc); p.nnzpb.swap(nn); p.type.swap(nt);
    p.bia_nnz.resize(p.nnzb); p.bia_origin.resize(p.nnzb); p.bia_end.resize(p.nnzb);
    for (int i = 0; i < p.nnzb; ++i) {
        p.bia_nnz[i] = a[i].blk_nnz; p.bia_origin[i] = a[i].origin_idx; p.bia_end[i] = a[i].end_idx;
    }
}

// ---- row-owned segments (coo2block.h:510-544) -------------------------------
void build_segments(Plan& p) {
    const int W = p.o.waves;
    p.row_uniform.assign(p.gb_m, 1);
    std::vector<uint8_t> has(p.gb_m, 0);
    for (int i = 0; i < p.nnzb; ++i) {
        const int r = p.brow[i];
        has[r] = 1;
        p.row_uniform[r] = p.row_uniform[r] && (p.type[i] & kUniform);
    }
    p.seg_rows.clear(); p.seg_begin.clear(); p.seg_end.clear();
    p.seg_direct.clear(); p.seg_uniform.clear();
    for (int r = 0; r < p.gb_m; ++r) {
        const int b = p.row_ptr[r], e = p.row_ptr[r + 1];
        if (e == b) continue;
        const bool uni = has[r] && p.row_uniform[r];
        const bool direct = p.nnz < 10000 || uni;
        const int count = direct ? 1 : (e - b + W - 1) / W;
        for (int s = 0; s < count; ++s) {
            const int sb = direct ? b : b + s * W;
            p.seg_rows.push_back(r);
            p.seg_begin.push_back(sb);
            p.seg_end.push_back(direct ? e : std::min(sb + W, e));
            p.seg_direct.push_back(direct);
            p.seg_uniform.push_back(uni);
        }
    }
}

// ---- flat CSR + row patterns (coo2block.h:546-602) --------------------------
// Patterns are numbered by the first row that shows them; rows are grouped by
// their sorted column multiset.  Rows are hashed in parallel, then numbered in
// one ordered pass that compares candidate column lists exactly.
void build_patterns(Plan& p, const int* row, const int* col) {
    const int m = p.m;
    p.flat_row_ptr.assign(m + 1, 0);
    for (int i = 0; i < p.nnz; ++i) p.flat_row_ptr[row[i] + 1]++;
    for (int r = 0; r < m; ++r) p.flat_row_ptr[r + 1] += p.flat_row_ptr[r];
    p.flat_col.assign(col, col + p.nnz);   // row-major input: already grouped by row

    std::vector<int> sorted(p.flat_col);
    std::vector<uint64_t> h(m);
#pragma omp parallel for schedule(dynamic, 1024)
    for (int r = 0; r < m; ++r) {
        int* b = sorted.data() + p.flat_row_ptr[r];
        int* e = sorted.data() + p.flat_row_ptr[r + 1];
        std::sort(b, e);
        uint64_t x = mix64((uint64_t)(e - b));
        for (int* q = b; q < e; ++q) x = mix64(x ^ (uint64_t)(uint32_t)*q);
        h[r] = x;
    }
    auto same = [&](int r1, int r2) {
        const int l1 = p.flat_row_ptr[r1 + 1] - p.flat_row_ptr[r1];
        const int l2 = p.flat_row_ptr[r2 + 1] - p.flat_row_ptr[r2];
        return l1 == l2 && std::equal(sorted.begin() + p.flat_row_ptr[r1],
                                      sorted.begin() + p.flat_row_ptr[r1 + 1],
                                      sorted.begin() + p.flat_row_ptr[r2]);
    };
    std::unordered_multimap<uint64_t, int> first;   // hash -> pattern id
    first.reserve(m);
    std::vector<int> pat_first_row;                   // pattern id -> its first row
    std::vector<int> row_pat(m);
    for (int r = 0; r < m; ++r) {
        int id = -1;
        auto range = first.equal_range(h[r]);
        for (auto it = range.first; it != range.second; ++it)
            if (same(pat_first_row[it->second], r)) { id = it->second; break; }
        if (id < 0) {
            id = (int)pat_first_row.size();
            pat_first_row.push_back(r);
            first.emplace(h[r], id);
        }
        row_pat[r] = id;
    }
    const int P = (int)pat_first_row.size();
    p.pat_row_ptr.assign(P + 1, 0);
    p.pat_out_ptr.assign(P + 1, 0);
    for (int q = 0; q < P; ++q) {
        const int r = pat_first_row[q];
        p.pat_row_ptr[q + 1] = p.pat_row_ptr[q] + (p.flat_row_ptr[r + 1] - p.flat_row_ptr[r]);
    }
    for (int r = 0; r < m; ++r) p.pat_out_ptr[row_pat[r] + 1]++;
    for (int q = 0; q < P; ++q) p.pat_out_ptr[q + 1] += p.pat_out_ptr[q];
    p.pat_col.assign(p.pat_row_ptr[P], 0);
#pragma omp parallel for schedule(dynamic, 256)
    for (int q = 0; q < P; ++q) {
        const int r = pat_first_row[q];
        std::copy(sorted.begin() + p.flat_row_ptr[r], sorted.begin() + p.flat_row_ptr[r + 1],
                  p.pat_col.begin() + p.pat_row_ptr[q]);
    }
    p.pat_out_rows.assign(m, 0);
    std::vector<int> cur(p.pat_out_ptr.begin(), p.pat_out_ptr.end() - 1);
    for (int r = 0; r < m; ++r) p.pat_out_rows[cur[row_pat[r]]++] = r;
}

}  // namespace

int build_plan(Plan& p, int m, int n, int nnz, const int* row, const int* col,
               const double* val, std::string& err) {
    const auto t0 = std::chrono::steady_clock::now();
    const Options& o = p.o;
    const int bs = o.bs, bcb = coord_bits(bs);
    p.m = m; p.n = n; p.nnz = nnz;
    p.gb_m = (m + bs - 1) / bs;
    p.gb_n = (n + bs - 1) / bs;

    // Validate the COO stream contract (a row-major stream, coo2block.h:500-508
    // and io/mmio_highlevel.h:392-398) and index ranges.
    for (int i = 0; i < nnz; ++i) {
        if (row[i] < 0 || row[i] >= m || col[i] < 0 || col[i] >= n) {
            err = "COO index out of range at entry " + std::to_string(i);
            return CBSPMV_ERR_RANGE;
        }
        if (i > 0 && row[i] < row[i - 1]) {
            err = "COO rows are not row-major (non-decreasing) at entry " + std::to_string(i);
            return CBSPMV_ERR_ORDER;
        }
    }
    p.in_row.assign(row, row + nnz);
    p.in_col.assign(col, col + nnz);
    p.in_val.assign(val, val + nnz);

    // entry range of each block row
    std::vector<int64_t> br_ptr(p.gb_m + 1, 0);
    for (int i = 0; i < nnz; ++i) br_ptr[row[i] / bs + 1]++;
    for (int b = 0; b < p.gb_m; ++b) br_ptr[b + 1] += br_ptr[b];

    const int nthreads = omp_get_max_threads();
    std::vector<Numbering> scratch(nthreads);

    // ---- pass 1: raw block statistics in original column space (gather_flag,
    // coo2block.h:63-75).  COMP builds skip the decision (never gather).
    bool gather = false;
    if (o.builder == CBSPMV_BUILDER_GATHER) {
        long long blocks = 0, small = 0;
#pragma omp parallel reduction(+ : blocks, small)
        {
            Numbering& nb = scratch[omp_get_thread_num()];
            nb.ensure(p.gb_n);
            std::vector<int> counts;
#pragma omp for schedule(dynamic, 64)
            for (int b = 0; b < p.gb_m; ++b) {
                counts.clear();
                for (int64_t i = br_ptr[b]; i < br_ptr[b + 1]; ++i) {
                    const int bj = col[i] / bs;
                    if (nb.stamp[bj] != b) { nb.stamp[bj] = b; nb.id[bj] = (int)counts.size(); counts.push_back(0); }
                    counts[nb.id[bj]]++;
                }
                blocks += (long long)counts.size();
                for (int c : counts) small += c <= 8;
            }
        }
        p.super_sparse = (int)small;
        p.analyzed = (int)blocks;
        p.super_ratio = blocks > 0 ? (double)small * 1.0 / (double)blocks : 0.0;
        gather = nnz >= 10 && p.super_ratio > 0.15;
        for (auto& s : scratch) std::fill(s.stamp.begin(), s.stamp.end(), -1);
    }
    p.gather = gather;

    // ---- pass 2: gathered columns (coo2block.h:7-61), first-appearance block
    // numbering per block row in the working column space, entries grouped by
    // block in input order (coo2block_pre_payload, :123-146).
    std::vector<int> wcol(nnz);               // working column (gathered or original)
    std::vector<int> restore_count(p.gb_m, 0);
    std::vector<std::vector<int>> restore(gather ? p.gb_m : 0);
    std::vector<int> local_blk(nnz);          // block id local to its block row
    std::vector<int> br_blocks(p.gb_m, 0);
#pragma omp parallel
    {
        Numbering& nb = scratch[omp_get_thread_num()];
        nb.ensure(p.gb_n);
#pragma omp for schedule(dynamic, 64)
        for (int b = 0; b < p.gb_m; ++b) {
            const int64_t s = br_ptr[b], e = br_ptr[b + 1];
            if (gather) {
                std::vector<int>& list = restore[b];
                list.assign(col + s, col + e);
                std::sort(list.begin(), list.end());
                list.erase(std::unique(list.begin(), list.end()), list.end());
                for (int64_t i = s; i < e; ++i)
                    wcol[i] = (int)(std::lower_bound(list.begin(), list.end(), col[i]) - list.begin());
                restore_count[b] = (int)list.size();
            } else {
                for (int64_t i = s; i < e; ++i) wcol[i] = col[i];
            }
            int k = 0;
            for (int64_t i = s; i < e; ++i) {
                const int bj = wcol[i] / bs;
                if (nb.stamp[bj] != b) { nb.stamp[bj] = b; nb.id[bj] = k++; }
                local_blk[i] = nb.id[bj];
            }
            br_blocks[b] = k;
        }
    }
    if (gather) {
        p.block_cols_off.assign(p.gb_m + 1, 0);
        for (int b = 0; b < p.gb_m; ++b) p.block_cols_off[b + 1] = p.block_cols_off[b] + restore_count[b];
        p.block_cols.resize(p.block_cols_off[p.gb_m]);
#pragma omp parallel for schedule(dynamic, 256)
        for (int b = 0; b < p.gb_m; ++b)
            std::copy(restore[b].begin(), restore[b].end(), p.block_cols.begin() + p.block_cols_off[b]);
        std::vector<std::vector<int>>().swap(restore);
    }
    std::vector<int> br_first(p.gb_m + 1, 0);   // global id of a block row's first block
    for (int b = 0; b < p.gb_m; ++b) br_first[b + 1] = br_first[b] + br_blocks[b];
    const int nnzb = br_first[p.gb_m];
    p.nnzb = nnzb;
    p.brow.assign(nnzb, 0); p.bcol.assign(nnzb, 0); p.nnzpb.assign(nnzb, 0);
    p.type.assign(nnzb, kCOO);
    p.raw_nnz.assign(nnzb, 0);
    p.pl_start.assign(nnzb + 1, 0);
    std::vector<int> perm(nnz);               // entries grouped by block, input order inside
#pragma omp parallel for schedule(dynamic, 64)
    for (int b = 0; b < p.gb_m; ++b) {
        const int64_t s = br_ptr[b], e = br_ptr[b + 1];
        const int g0 = br_first[b], k = br_blocks[b];
        for (int64_t i = s; i < e; ++i) {
            const int g = g0 + local_blk[i];
            if (p.raw_nnz[g]++ == 0) { p.brow[g] = b; p.bcol[g] = wcol[i] / bs; }
        }
        // block payload slots are laid out in block order inside the row slice
        int64_t at = s;
        for (int q = 0; q < k; ++q) { p.pl_start[g0 + q] = at; at += p.raw_nnz[g0 + q]; }
        std::vector<int64_t> cur(k);
        for (int q = 0; q < k; ++q) cur[q] = p.pl_start[g0 + q];
        for (int64_t i = s; i < e; ++i) perm[cur[local_blk[i]]++] = (int)i;
    }
    p.pl_start[nnzb] = nnz;
    for (int g = 0; g < nnzb; ++g) p.nnzpb[g] = p.raw_nnz[g];
    std::vector<int>().swap(local_blk);

    // ---- pass 3: format selection and encoding (coo2block_post, :154-380)
    const int dense_thr = bs * bs * o.dense_sixteenths / 16;
    p.direct = gather && n <= 65535;
    p.compact = p.direct && bs == 16 && n <= 4095 && nnz < 10000;
    long long cand = 0, bal = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(+ : cand, bal)
    for (int g = 0; g < nnzb; ++g) {
        const int k = p.raw_nnz[g];
        if (k <= 32 || k > 64) continue;
        int cnt[64] = {0};
        for (int64_t q = p.pl_start[g]; q < p.pl_start[g] + k; ++q)
            cnt[trunc_orig(bs, row[perm[q]] - p.brow[g] * bs)]++;
        int lo = INT_MAX, hi = 0;
        for (int r = 0; r < bs; ++r)
            if (cnt[r] > 0) { lo = std::min(lo, cnt[r]); hi = std::max(hi, cnt[r]); }
        ++cand;
        bal += (hi - lo <= 1);
    }
    const bool extend = o.force_coo64 || (cand > 0 && 2 * bal < cand);
    auto is_coo = [extend](int k) { return k <= 32 || (extend && k <= 64); };

    p.dense_slot.assign(nnzb, -1);
    p.csr_rp_start.assign(nnzb, -1);
    int ndense = 0, ncsr = 0;
    for (int g = 0; g < nnzb; ++g) {
        const int k = p.raw_nnz[g];
        if (is_coo(k)) continue;
        if (k >= dense_thr) p.dense_slot[g] = ndense++;
        else p.csr_rp_start[g] = (int64_t)(ncsr++) * (bs + 1);
    }
    p.coo_idx.assign(nnz, 0);
    p.ent_val.assign(nnz, 0.0);
    p.csr_col.assign(nnz, 0);
    p.csr_rp.assign((size_t)ncsr * (bs + 1), 0);
    p.dense_val.assign((size_t)ndense * bs * bs, 0.0);
    p.coo_count.assign(nnzb, 0);

    const bool direct = p.direct, compact = p.compact;
    const bool position_bits = bs != 64 && ((direct && !compact) || bs == 32);
    const int rshift = direct ? 16 : bcb;
    const int pshift = direct ? 16 + bcb : 10;
    const uint32_t longflag = direct ? (1u << (pshift + 5)) : 0x8000u;
    const uint32_t rmask = (1u << bcb) - 1;

#pragma omp parallel
    {
        std::vector<int> lr, lc, ord;
        std::vector<double> lv, dense(bs * bs);
#pragma omp for schedule(dynamic, 256)
        for (int g = 0; g < nnzb; ++g) {
            const int k = p.raw_nnz[g];
            const int64_t s = p.pl_start[g];
            lr.resize(k); lc.resize(k); lv.resize(k);
            for (int q = 0; q < k; ++q) {
                const int i = perm[s + q];
                lr[q] = trunc_orig(bs, row[i] - p.brow[g] * bs);
                lc[q] = trunc_orig(bs, wcol[i] - p.bcol[g] * bs);
                lv[q] = val[i];
            }
            auto uniform_of = [&]() {
                for (int r = 1; r < bs; ++r)
                    for (int c = 0; c < bs; ++c)
                        if (!(dense[r * bs + c] == dense[c])) return false;
                return true;
            };
            if (is_coo(k)) {
                ord.resize(k);
                std::iota(ord.begin(), ord.end(), 0);
                std::sort(ord.begin(), ord.end(), [&](int a, int b2) {
                    if (lr[a] != lr[b2]) return lr[a] < lr[b2];
                    return lc[a] < lc[b2];
                });
                uint32_t* idx = p.coo_idx.data() + s;
                double* vv = p.ent_val.data() + s;
                int kk = 0;
                for (int j : ord) {
                    const uint32_t r = (uint32_t)lr[j], c = (uint32_t)lc[j];
                    uint32_t w;
                    if (direct) {
                        const uint32_t oc = (uint32_t)p.block_cols[p.block_cols_off[p.brow[g]] + p.bcol[g] * bs + (int)c];
                        w = compact ? ((r << 12) | oc) : ((r << 16) | oc);
                    } else {
                        w = (r << bcb) | c;
                    }
                    if (kk > 0 && idx[kk - 1] == w) vv[kk - 1] += lv[j];
                    else { idx[kk] = w; vv[kk] = lv[j]; ++kk; }
                }
                if (position_bits) {
                    bool long_rows = false;
                    int a = 0;
                    while (a < kk) {
                        const uint32_t rr = (idx[a] >> rshift) & rmask;
                        int e = a + 1;
                        while (e < kk && ((idx[e] >> rshift) & rmask) == rr) ++e;
                        for (int q = 0; a + q < e; ++q) idx[a + q] |= (uint32_t)(q << pshift);
                        long_rows = long_rows || e - a > 8;
                        a = e;
                    }
                    if (long_rows) for (int q = 0; q < kk; ++q) idx[q] |= longflag;
                }
                std::fill(dense.begin(), dense.end(), 0.0);
                for (int q = 0; q < k; ++q) dense[lr[q] * bs + lc[q]] += lv[q];
                p.type[g] = kCOO | (uniform_of() ? kUniform : 0);
                p.nnzpb[g] = kk;
                p.coo_count[g] = kk;
            } else if (k >= dense_thr) {
                std::fill(dense.begin(), dense.end(), 0.0);
                for (int q = 0; q < k; ++q) dense[lr[q] * bs + lc[q]] += lv[q];
                p.type[g] = kDENSE | (uniform_of() ? kUniform : 0);
                std::copy(dense.begin(), dense.end(), p.dense_val.begin() + (size_t)p.dense_slot[g] * bs * bs);
            } else {
                int* rp = p.csr_rp.data() + p.csr_rp_start[g];
                std::fill(rp, rp + bs + 1, 0);
                for (int q = 0; q < k; ++q) rp[lr[q]] = trunc_orig(bs, rp[lr[q]] + 1);
                for (int r = 0, run = 0; r < bs; ++r) { const int t = rp[r]; rp[r] = trunc_orig(bs, run); run += t; }
                rp[bs] = trunc_orig(bs, k);
                int cur[65];
                for (int r = 0; r <= bs; ++r) cur[r] = rp[r];
                int* cc = p.csr_col.data() + s;
                double* cv = p.ent_val.data() + s;
                for (int q = 0; q < k; ++q) {
                    const int dst = cur[lr[q]];
                    cur[lr[q]] = trunc_orig(bs, cur[lr[q]] + 1);
                    cc[dst] = lc[q];
                    cv[dst] = lv[q];
                }
                bool uni = true;
                const int len0 = rp[1] - rp[0];
                for (int r = 1; r < bs && uni; ++r) {
                    const int b0 = rp[r], len = rp[r + 1] - b0;
                    if (len != len0) { uni = false; break; }
                    for (int q = 0; q < len; ++q)
                        uni = uni && cc[b0 + q] == cc[q] && cv[b0 + q] == cv[q];
                }
                p.type[g] = kCSR | (uni ? kUniform : 0);
            }
        }
    }
    std::vector<int>().swap(perm);
    std::vector<int>().swap(wcol);

    // ---- row pointer, segments or LPT balance, patterns (:652-669)
    p.row_ptr.assign(p.gb_m + 1, 0);
    for (int g = 0; g < nnzb; ++g) p.row_ptr[p.brow[g] + 1]++;
    for (int b = 0; b < p.gb_m; ++b) p.row_ptr[b + 1] += p.row_ptr[b];
    if (o.row_owned) build_segments(p);
    if (!o.row_owned && nnz >= 10000) {
        balance(p);
    } else {
        p.bia_nnz = p.nnzpb;
        p.bia_origin.resize(p.nnzb);
        p.bia_end.resize(p.nnzb);
        std::iota(p.bia_origin.begin(), p.bia_origin.end(), 0);
        std::iota(p.bia_end.begin(), p.bia_end.end(), 0);
    }
    build_patterns(p, row, col);

    p.build_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return CBSPMV_OK;
}

}  // namespace cbs
