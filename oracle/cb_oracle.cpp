// CB-SpMV oracle: a plain, single-threaded CPU restatement of the reference
// algorithm, used ONLY as the oracle by tests/, __graft_entry__.smoke() and
// the cpu_baseline leg of bench.py.  Nothing in the product links or loads it.
//
// What it restates (reference paths are relative to the reference checkout):
//   * sequential FP64 SpMV            cb-spmv-dcu/src/utils.h:98-111
//   * the CB builder                  cb-spmv-dcu/src/coo2block.h:7-670
//   * the packed-slab layout          cb-spmv-dcu/src/cb-spmv-dcu.hpp:1100-1348
//   * the dispatch policy             cb-spmv-dcu/src/cb-spmv-dcu.hpp:1105-1127, 1428-1450
//
// It is C++ rather than C for one reason: the reference breaks ties of two
// unstable std::sort calls (COO entry order inside a block, coo2block.h:219-227,
// and the LPT block order, coo2block.h:409) in whatever order libstdc++'s
// introsort leaves them.  Duplicate (row, col) entries and blocks of equal nnz
// are therefore ordered by calling std::sort with an equivalent comparator on
// an identical input sequence.  Everything else is plain loops over arrays.
//
// How each part is pinned (tests/test_golden.py, test_builder_parity.py,
// test_oracle.py):
//   * builder arrays and payloads: equal to the reference builder compiled from
//     its own sources (oracle/_ref), on the committed fixtures and on seeded
//     random inputs; plus closed-form checks (floor-division block ids,
//     numpy.unique restore maps, thresholds at 32/33/64/65/239/240 nnz, the
//     n = 4095/4096/65535/65536 and nnz = 9999/10000 switch points).
//   * SpMV: scipy.sparse CSR product, e_j column probes and linearity.
//   * layout: alignment/disjointness/ordering invariants and the kernels
//     running on it.  The reference packer itself allocates device memory
//     before it computes addresses, so it cannot run here: the byte layout is
//     "parity pinned by invariants", not by a reference image.
#include <omp.h>
#include <sched.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <thread>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <queue>
#include <string>
#include <unordered_map>
#include <vector>

namespace {

struct Opts {
    int bs;                // block size: 2, 4, 8, 16, 32 or 64   (macros.h:38-74)
    int waves;             // waves per workgroup                   (macros.h:80-82)
    int row_owned;         // row-owned segments (1) or LPT (0)     (macros.h:86-88)
    int force_coo64;       //                                       (macros.h:89-91)
    int dense_sixteenths;  //                                       (macros.h:92-94)
    int pattern_wave;      // pattern arrays packed into the slab   (macros.h:95-97)
};

enum : uint8_t { kCOO = 0, kCSR = 1, kDENSE = 3, kTypeMask = 0x7f, kUniform = 0x80 };

int coord_bits(int bs) { int b = 0; while ((1 << b) < bs) ++b; return b; }
size_t orig_index_bytes(int bs) { return bs <= 16 ? 1 : 2; }           // IndexTypeOrig
size_t dense_mask_bytes(int bs) { return bs <= 8 ? 1 : bs == 16 ? 2 : bs == 32 ? 4 : 8; }
size_t align_up(size_t v, size_t a) { return (v + a - 1) & ~(a - 1); }

// IndexTypeOrig truncation (uint8 for BS<=16, uint16 above).
int orig_trunc(int bs, int v) { return bs <= 16 ? (v & 0xff) : (v & 0xffff); }

struct OrigBlock {            // COO_BLOCK_ORIG, entries in input order
    std::vector<int> r, c;    // local coordinates (already IndexTypeOrig-truncated)
    std::vector<double> v;
};

struct Plan {
    Opts o{};
    int m = 0, n = 0, nnz = 0;
    int gb_m = 0, gb_n = 0, nnzb = 0;
    bool gather = false, direct = false, compact = false;
    int super_sparse = 0, analyzed = 0;
    double super_ratio = 0.0;
    bool balanced = false;
    double sd_before = 0.0, sd_after = 0.0;

    std::vector<int> brow, bcol, nnzpb, row_ptr;
    std::vector<uint8_t> type;
    std::vector<int> seg_rows, seg_begin, seg_end;
    std::vector<uint8_t> seg_direct, seg_uniform, row_uniform;
    std::vector<int> flat_row_ptr, flat_col;
    std::vector<int> pat_row_ptr, pat_col, pat_out_ptr, pat_out_rows;
    std::vector<int> block_cols, block_cols_off;
    std::vector<int> bia_nnz, bia_origin, bia_end;   // BLOCK_IDX array

    // payloads keyed by the pre-balance block index ("origin index")
    std::vector<OrigBlock> orig;
    std::vector<std::vector<uint32_t>> coo_idx;
    std::vector<std::vector<double>> coo_val;
    std::vector<std::vector<int>> csr_rp, csr_col;
    std::vector<std::vector<double>> csr_val;
    std::vector<std::vector<double>> dense_val;

    // packed image (computed on demand)
    std::vector<uint8_t> image;
    std::map<std::string, size_t> off;
    bool unit_values = false;
};

// ---- coo2block_pre_metadata (coo2block.h:77-121): first-appearance block ids
void pre_metadata(Plan& p, const int* row, const int* col, std::vector<int>& ids) {
    const int bs = p.o.bs;
    ids.assign(p.nnz, 0);
    std::unordered_map<long long, int> seen;
    for (int i = 0; i < p.nnz; ++i) {
        const int bi = row[i] / bs, bj = col[i] / bs;
        const long long key = (long long)bi * p.gb_n + bj;
        auto it = seen.find(key);
        if (it == seen.end()) {
            const int id = p.nnzb++;
            seen.emplace(key, id);
            p.brow.push_back(bi);
            p.bcol.push_back(bj);
            p.nnzpb.push_back(1);
            p.type.push_back(kCOO);
            ids[i] = id;
        } else {
            p.nnzpb[it->second]++;
            ids[i] = it->second;
        }
    }
}

// ---- coo2block_pre_payload (coo2block.h:123-146): scatter in input order
void pre_payload(Plan& p, const int* row, const int* col, const double* val,
                 const std::vector<int>& ids) {
    const int bs = p.o.bs;
    p.orig.assign(p.nnzb, OrigBlock{});
    for (int i = 0; i < p.nnz; ++i) {
        OrigBlock& b = p.orig[ids[i]];
        b.r.push_back(orig_trunc(bs, row[i] - p.brow[ids[i]] * bs));
        b.c.push_back(orig_trunc(bs, col[i] - p.bcol[ids[i]] * bs));
        b.v.push_back(val[i]);
    }
}

// ---- gather_flag (coo2block.h:63-75)
bool gather_flag(Plan& p) {
    int small = 0;
    for (int i = 0; i < p.nnzb; ++i) small += p.nnzpb[i] <= 8;
    p.super_sparse = small;
    p.analyzed = p.nnzb;
    p.super_ratio = p.nnzb > 0 ? small * 1.0 / p.nnzb : 0.0;
    return p.nnz >= 10 && p.super_ratio > 0.15;
}

// ---- gather (coo2block.h:7-61): per block row, sorted unique original
// columns; gathered column = position in that list.
void gather(Plan& p, const int* row, const int* col, std::vector<int>& gcol) {
    const int bs = p.o.bs;
    std::vector<std::vector<int>> cols(p.gb_m);
    for (int i = 0; i < p.nnz; ++i) cols[row[i] / bs].push_back(col[i]);
    for (auto& c : cols) {
        std::sort(c.begin(), c.end());
        c.erase(std::unique(c.begin(), c.end()), c.end());
    }
    gcol.assign(p.nnz, 0);
    for (int i = 0; i < p.nnz; ++i) {
        const auto& c = cols[row[i] / bs];
        gcol[i] = (int)(std::lower_bound(c.begin(), c.end(), col[i]) - c.begin());
    }
    p.block_cols_off.assign(p.gb_m + 1, 0);
    for (int i = 0; i < p.gb_m; ++i) p.block_cols_off[i + 1] = p.block_cols_off[i] + (int)cols[i].size();
    p.block_cols.clear();
    for (auto& c : cols) p.block_cols.insert(p.block_cols.end(), c.begin(), c.end());
}

// ---- coo2block_post (coo2block.h:154-380): format choice and encodings
void post(Plan& p) {
    const int bs = p.o.bs, bcb = coord_bits(bs);
    const int dense_thr = bs * bs * p.o.dense_sixteenths / 16;
    p.direct = p.gather && p.n <= 65535;
    p.compact = p.direct && bs == 16 && p.n <= 4095 && p.nnz < 10000;

    // COO-64 extension vote (coo2block.h:162-185)
    int cand = 0, bal = 0;
    for (const OrigBlock& b : p.orig) {
        const int k = (int)b.v.size();
        if (k <= 32 || k > 64) continue;
        std::vector<int> cnt(bs, 0);
        for (int r : b.r) cnt[r]++;
        int lo = INT_MAX, hi = 0;
        for (int r = 0; r < bs; ++r)
            if (cnt[r] > 0) { lo = std::min(lo, cnt[r]); hi = std::max(hi, cnt[r]); }
        ++cand;
        bal += (hi - lo <= 1);
    }
    const bool extend = p.o.force_coo64 || (cand > 0 && 2 * bal < cand);

    p.coo_idx.assign(p.nnzb, {});
    p.coo_val.assign(p.nnzb, {});
    p.csr_rp.assign(p.nnzb, {});
    p.csr_col.assign(p.nnzb, {});
    p.csr_val.assign(p.nnzb, {});
    p.dense_val.assign(p.nnzb, {});

    auto rows_uniform_dense = [bs](const std::vector<double>& d) {
        for (int r = 1; r < bs; ++r)
            for (int c = 0; c < bs; ++c)
                if (!(d[r * bs + c] == d[c])) return false;
        return true;
    };

    for (int i = 0; i < p.nnzb; ++i) {
        const OrigBlock& b = p.orig[i];
        const int k = (int)b.v.size();
        const bool is_coo = k <= 32 || (extend && k <= 64);
        if (is_coo) {
            // entry order: (row, col) ascending with std::sort (ties: duplicates)
            std::vector<int> ord(k);
            for (int j = 0; j < k; ++j) ord[j] = j;
            std::sort(ord.begin(), ord.end(), [&b](int a, int c) {
                if (b.r[a] != b.r[c]) return b.r[a] < b.r[c];
                return b.c[a] < b.c[c];
            });
            std::vector<uint32_t>& idx = p.coo_idx[i];
            std::vector<double>& vv = p.coo_val[i];
            for (int j : ord) {
                const uint32_t lr = (uint32_t)b.r[j], lc = (uint32_t)b.c[j];
                uint32_t w;
                if (p.direct) {
                    const int gc = p.bcol[i] * bs + (int)lc;
                    const uint32_t oc = (uint32_t)p.block_cols[p.block_cols_off[p.brow[i]] + gc];
                    w = p.compact ? ((lr << 12) | oc) : ((lr << 16) | oc);
                } else {
                    w = (lr << bcb) | lc;
                }
                if (!idx.empty() && idx.back() == w) {
                    vv.back() += b.v[j];       // merge duplicates in sorted order
                } else {
                    idx.push_back(w);
                    vv.push_back(b.v[j]);
                }
            }
            const int kk = (int)idx.size();
            // in-row position and long-row flag (coo2block.h:258-291)
            if (bs != 64 && ((p.direct && !p.compact) || bs == 32)) {
                const int rshift = p.direct ? 16 : bcb;
                const int pshift = p.direct ? 16 + bcb : 10;
                const uint32_t longflag = p.direct ? (1u << (pshift + 5)) : 0x8000u;
                const uint32_t rmask = (1u << bcb) - 1;
                bool long_rows = false;
                int s = 0;
                while (s < kk) {
                    const uint32_t rr = (idx[s] >> rshift) & rmask;
                    int e = s + 1;
                    while (e < kk && ((idx[e] >> rshift) & rmask) == rr) ++e;
                    for (int q = 0; s + q < e; ++q) idx[s + q] |= (uint32_t)(q << pshift);
                    if (e - s > 8) long_rows = true;
                    s = e;
                }
                if (long_rows)
                    for (uint32_t& w : idx) w |= longflag;
            }
            // uniform-rows test on the input-order dense image (coo2block.h:292-307)
            std::vector<double> d(bs * bs, 0.0);
            for (int j = 0; j < k; ++j) d[b.r[j] * bs + b.c[j]] += b.v[j];
            p.type[i] = kCOO | (rows_uniform_dense(d) ? kUniform : 0);
            p.nnzpb[i] = kk;
        } else if (k >= dense_thr) {
            std::vector<double> d(bs * bs, 0.0);
            for (int j = 0; j < k; ++j) d[b.r[j] * bs + b.c[j]] += b.v[j];
            p.type[i] = kDENSE | (rows_uniform_dense(d) ? kUniform : 0);
            p.dense_val[i] = d;
        } else {
            // CSR: stable by row, input order inside a row (coo2block.h:331-377)
            std::vector<int> rp(bs + 1, 0);
            for (int j = 0; j < k; ++j) rp[b.r[j]]++;
            int run = 0;
            for (int r = 0; r < bs; ++r) { const int t = rp[r]; rp[r] = orig_trunc(bs, run); run += t; }
            rp[bs] = orig_trunc(bs, k);
            std::vector<int> cur(rp.begin(), rp.end());
            std::vector<int> cc(k, 0);
            std::vector<double> cv(k, 0.0);
            for (int j = 0; j < k; ++j) {
                const int dst = cur[b.r[j]];
                cur[b.r[j]] = orig_trunc(bs, cur[b.r[j]] + 1);
                cc[dst] = b.c[j];
                cv[dst] = b.v[j];
            }
            bool uni = true;
            const int len0 = rp[1] - rp[0];
            for (int r = 1; r < bs && uni; ++r) {
                const int s0 = rp[r], len = rp[r + 1] - s0;
                if (len != len0) { uni = false; break; }
                for (int q = 0; q < len; ++q)
                    uni = uni && cc[s0 + q] == cc[q] && cv[s0 + q] == cv[q];
            }
            p.type[i] = kCSR | (uni ? kUniform : 0);
            p.csr_rp[i] = rp;
            p.csr_col[i] = cc;
            p.csr_val[i] = cv;
        }
    }
}

// ---- coo2block_build_row_ptr (coo2block.h:500-508)
void build_row_ptr(Plan& p) {
    p.row_ptr.assign(p.gb_m + 1, 0);
    for (int i = 0; i < p.nnzb; ++i) p.row_ptr[p.brow[i] + 1]++;
    for (int r = 0; r < p.gb_m; ++r) p.row_ptr[r + 1] += p.row_ptr[r];
}

// ---- coo2block_build_segments (coo2block.h:510-544)
void build_segments(Plan& p) {
    const int W = p.o.waves;
    p.row_uniform.assign(p.gb_m, 1);
    std::vector<uint8_t> has(p.gb_m, 0);
    for (int i = 0; i < p.nnzb; ++i) {
        has[p.brow[i]] = 1;
        p.row_uniform[p.brow[i]] = p.row_uniform[p.brow[i]] && (p.type[i] & kUniform);
    }
    for (int r = 0; r < p.gb_m; ++r) {
        const int b = p.row_ptr[r], e = p.row_ptr[r + 1];
        const bool uni = has[r] && p.row_uniform[r];
        const bool direct = p.nnz < 10000 || uni;
        const int count = (e == b) ? 0 : direct ? 1 : (e - b + W - 1) / W;
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

double group_stddev(const std::vector<int>& loads) {
    double sum = 0.0;
    for (int l : loads) sum += l;
    const double mean = sum / loads.size();
    double var = 0.0;
    for (int l : loads) { const double d = l - mean; var += d * d; }
    return std::sqrt(var / loads.size());
}

// ---- coo2block_banlance (coo2block.h:382-498): LPT over groups of W blocks
struct Bia { int nnz, origin, end; };

void balance(Plan& p) {
    const int W = p.o.waves;
    std::vector<Bia> a(p.nnzb);
    for (int i = 0; i < p.nnzb; ++i) a[i] = {p.nnzpb[i], i, 0};
    std::vector<int> before((p.nnzb + W - 1) / W, 0);
    for (int i = 0; i < p.nnzb; ++i) before[i / W] += p.nnzpb[i];
    const double sdb = group_stddev(before);

    std::sort(a.begin(), a.end(), [](Bia x, Bia y) { return x.nnz > y.nnz; });
    int groups;
    if (p.nnzb % W != 0) {
        groups = p.nnzb / W + 1;
        const int pad = W - p.nnzb % W;
        for (int q = 0; q < pad; ++q) a.push_back({0, INT_MAX, -1});
        p.nnzb += pad;
    } else {
        groups = p.nnzb / W;
    }
    struct Item { int total, index, count; };
    auto worse = [](const Item& x, const Item& y) {   // min-heap on (total, index)
        if (x.total == y.total) return x.index > y.index;
        return x.total > y.total;
    };
    std::priority_queue<Item, std::vector<Item>, decltype(worse)> pq(worse);
    for (int g = 0; g < groups; ++g) pq.push({0, g, 0});
    for (int i = 0; i < p.nnzb; ++i) {
        Item t = pq.top();
        pq.pop();
        a[i].end = t.index * W + t.count;
        t.total += a[i].nnz;
        t.count += 1;
        if (t.count < W) pq.push(t);
    }
    std::sort(a.begin(), a.end(), [](Bia x, Bia y) { return x.end < y.end; });

    std::vector<int> nr(p.nnzb), nc(p.nnzb), nn(p.nnzb);
    std::vector<uint8_t> nt(p.nnzb);
    for (int i = 0; i < p.nnzb; ++i) {
        const int o = a[i].origin;
        if (o != INT_MAX) { nr[i] = p.brow[o]; nc[i] = p.bcol[o]; nn[i] = p.nnzpb[o]; nt[i] = p.type[o]; }
        else { nr[i] = -1; nc[i] = -1; nn[i] = 0; nt[i] = kCOO; }
    }
    std::vector<int> after(groups, 0);
    for (int i = 0; i < p.nnzb; ++i) after[i / W] += a[i].nnz;
    p.balanced = true;
    p.sd_before = sdb;
    p.sd_after = group_stddev(after);
    p.brow = nr; p.bcol = nc; p.nnzpb = nn; p.type = nt;
    p.bia_nnz.resize(p.nnzb); p.bia_origin.resize(p.nnzb); p.bia_end.resize(p.nnzb);
    for (int i = 0; i < p.nnzb; ++i) { p.bia_nnz[i] = a[i].nnz; p.bia_origin[i] = a[i].origin; p.bia_end[i] = a[i].end; }
}

void identity_bia(Plan& p) {
    p.bia_nnz = p.nnzpb;
    p.bia_origin.resize(p.nnzb);
    p.bia_end.resize(p.nnzb);
    for (int i = 0; i < p.nnzb; ++i) { p.bia_origin[i] = i; p.bia_end[i] = i; }
}

// ---- coo2block_build_flat_csr + coo2block_build_patterns (coo2block.h:546-602)
void build_patterns(Plan& p, const int* row, const int* col) {
    p.flat_row_ptr.assign(p.m + 1, 0);
    for (int i = 0; i < p.nnz; ++i) p.flat_row_ptr[row[i] + 1]++;
    for (int r = 0; r < p.m; ++r) p.flat_row_ptr[r + 1] += p.flat_row_ptr[r];
    p.flat_col.assign(p.nnz, 0);
    std::vector<int> cur(p.flat_row_ptr.begin(), p.flat_row_ptr.end() - 1);
    for (int i = 0; i < p.nnz; ++i) p.flat_col[cur[row[i]]++] = col[i];

    std::map<std::vector<int>, int> ids;
    std::vector<std::vector<int>> pcols, prows;
    for (int r = 0; r < p.m; ++r) {
        std::vector<int> c(p.flat_col.begin() + p.flat_row_ptr[r], p.flat_col.begin() + p.flat_row_ptr[r + 1]);
        std::sort(c.begin(), c.end());
        auto it = ids.find(c);
        int id;
        if (it == ids.end()) {
            id = (int)pcols.size();
            ids.emplace(c, id);
            pcols.push_back(c);
            prows.emplace_back();
        } else {
            id = it->second;
        }
        prows[id].push_back(r);
    }
    p.pat_row_ptr.assign(pcols.size() + 1, 0);
    p.pat_out_ptr.assign(pcols.size() + 1, 0);
    p.pat_col.clear();
    p.pat_out_rows.clear();
    for (size_t q = 0; q < pcols.size(); ++q) {
        p.pat_col.insert(p.pat_col.end(), pcols[q].begin(), pcols[q].end());
        p.pat_out_rows.insert(p.pat_out_rows.end(), prows[q].begin(), prows[q].end());
        p.pat_row_ptr[q + 1] = (int)p.pat_col.size();
        p.pat_out_ptr[q + 1] = (int)p.pat_out_rows.size();
    }
}

void finish(Plan& p, const int* row, const int* col) {
    build_row_ptr(p);
    if (p.o.row_owned) build_segments(p);
    if (!p.o.row_owned && p.nnz >= 10000) balance(p);
    else identity_bia(p);
    build_patterns(p, row, col);
}

void reset_blocks(Plan& p) {
    p.nnzb = 0;
    p.brow.clear(); p.bcol.clear(); p.nnzpb.clear(); p.type.clear();
}

// ---- coo2block_gather (coo2block.h:628-670)
Plan* build_gather(int m, int n, int nnz, const int* row, const int* col, const double* val, const Opts& o) {
    Plan* p = new Plan();
    p->o = o; p->m = m; p->n = n; p->nnz = nnz;
    p->gb_m = (m + o.bs - 1) / o.bs;
    p->gb_n = (n + o.bs - 1) / o.bs;
    std::vector<int> ids;
    pre_metadata(*p, row, col, ids);
    if (gather_flag(*p)) {
        std::vector<int> gcol;
        gather(*p, row, col, gcol);
        reset_blocks(*p);
        p->gather = true;
        pre_metadata(*p, row, gcol.data(), ids);
        pre_payload(*p, row, gcol.data(), val, ids);
    } else {
        pre_payload(*p, row, col, val, ids);
    }
    post(*p);
    finish(*p, row, col);
    return p;
}

// ---- coo2block_comp (coo2block.h:605-626): never gathers, no statistics
Plan* build_comp(int m, int n, int nnz, const int* row, const int* col, const double* val, const Opts& o) {
    Plan* p = new Plan();
    p->o = o; p->m = m; p->n = n; p->nnz = nnz;
    p->gb_m = (m + o.bs - 1) / o.bs;
    p->gb_n = (n + o.bs - 1) / o.bs;
    std::vector<int> ids;
    pre_metadata(*p, row, col, ids);
    pre_payload(*p, row, col, val, ids);
    post(*p);
    finish(*p, row, col);
    return p;
}

// ---- packed slab (cb-spmv-dcu.hpp:1100-1348).  Block addresses are stored
// as byte offsets from the slab base (0 = no payload) instead of pointers.
void pack(Plan& p) {
    const int bs = p.o.bs;
    bool unit = true;
    for (int i = 0; i < (int)p.coo_val.size(); ++i) for (double v : p.coo_val[i]) unit = unit && v == 1.0;
    for (int i = 0; i < (int)p.csr_val.size(); ++i) for (double v : p.csr_val[i]) unit = unit && v == 1.0;
    for (int i = 0; i < (int)p.dense_val.size(); ++i) for (double v : p.dense_val[i]) unit = unit && (v == 0.0 || v == 1.0);
    p.unit_values = unit;

    size_t cur = 0;
    auto take = [&cur](size_t bytes, size_t al) { cur = align_up(cur, al); const size_t o = cur; cur += bytes; return o; };
    const size_t nb = p.nnzb, S = p.seg_rows.size();
    auto& off = p.off;
    off.clear();
    off["gb_row_idx"] = take(nb * 4, 4);
    off["gb_col_idx"] = take(nb * 4, 4);
    if (p.o.row_owned) {
        off["segment_rows"] = take(S * 4, 4);
        off["segment_begin"] = take(S * 4, 4);
        off["segment_end"] = take(S * 4, 4);
        off["segment_direct"] = take(S, 1);
        off["segment_uniform"] = take(S, 1);
    }
    off["nnz_per_blk"] = take(nb * 4, 4);
    off["type_per_blk"] = take(nb, 1);
    if (p.o.pattern_wave) {
        off["pattern_row_ptr"] = take(p.pat_row_ptr.size() * 4, 4);
        off["pattern_col_idx"] = take(p.pat_col.size() * 4, 4);
        off["pattern_output_ptr"] = take(p.pat_out_ptr.size() * 4, 4);
        off["pattern_output_rows"] = take(p.pat_out_rows.size() * 4, 4);
    }
    if (p.gather) {
        off["block_cols"] = take(p.block_cols.size() * 4, 4);
        off["block_cols_offset"] = take(p.block_cols_off.size() * 4, 4);
    }
    off["block_addr"] = take(nb * 8, 8);
    std::vector<size_t> boff(nb, 0);
    const size_t ob = orig_index_bytes(bs);
    for (size_t i = 0; i < nb; ++i) {
        if (p.nnzpb[i] <= 0) continue;
        const int o = p.bia_origin[i];
        const int t = p.type[i] & kTypeMask;
        if (t == kCOO) {
            const size_t k = p.coo_idx[o].size();
            const size_t ib = k * (p.compact ? 2 : 4);
            boff[i] = take(unit ? ib : align_up(ib, 8) + k * 8, unit ? 4 : 8);
        } else if (t == kCSR) {
            const size_t k = p.csr_col[o].size();
            const size_t ib = (bs + 1) * ob + k * ob;
            boff[i] = take(unit ? ib : align_up(ib, 8) + k * 8, unit ? ob : 8);
        } else {
            boff[i] = take(unit ? bs * dense_mask_bytes(bs) : (size_t)bs * bs * 8, unit ? dense_mask_bytes(bs) : 8);
        }
    }
    off["x"] = take((size_t)p.gb_n * bs * 8, 8);
    off["y"] = take((size_t)p.gb_m * bs * 8, 8);
    if (p.o.row_owned) off["output_epochs"] = take((size_t)p.gb_m * 4, 4);
    off["total"] = cur;

    // the image is what the reference uploads: every byte before y
    p.image.assign(off["y"], 0);
    auto put = [&p](size_t o, const void* src, size_t bytes) { if (bytes) std::memcpy(p.image.data() + o, src, bytes); };
    put(off["gb_row_idx"], p.brow.data(), nb * 4);
    put(off["gb_col_idx"], p.bcol.data(), nb * 4);
    if (p.o.row_owned) {
        put(off["segment_rows"], p.seg_rows.data(), S * 4);
        put(off["segment_begin"], p.seg_begin.data(), S * 4);
        put(off["segment_end"], p.seg_end.data(), S * 4);
        put(off["segment_direct"], p.seg_direct.data(), S);
        put(off["segment_uniform"], p.seg_uniform.data(), S);
    }
    put(off["nnz_per_blk"], p.nnzpb.data(), nb * 4);
    put(off["type_per_blk"], p.type.data(), nb);
    if (p.o.pattern_wave) {
        put(off["pattern_row_ptr"], p.pat_row_ptr.data(), p.pat_row_ptr.size() * 4);
        put(off["pattern_col_idx"], p.pat_col.data(), p.pat_col.size() * 4);
        put(off["pattern_output_ptr"], p.pat_out_ptr.data(), p.pat_out_ptr.size() * 4);
        put(off["pattern_output_rows"], p.pat_out_rows.data(), p.pat_out_rows.size() * 4);
    }
    if (p.gather) {
        put(off["block_cols"], p.block_cols.data(), p.block_cols.size() * 4);
        put(off["block_cols_offset"], p.block_cols_off.data(), p.block_cols_off.size() * 4);
    }
    std::vector<uint64_t> addr(nb, 0);
    for (size_t i = 0; i < nb; ++i) if (p.nnzpb[i] > 0) addr[i] = boff[i];
    put(off["block_addr"], addr.data(), nb * 8);
    for (size_t i = 0; i < nb; ++i) {
        if (p.nnzpb[i] <= 0) continue;
        const int o = p.bia_origin[i];
        const int t = p.type[i] & kTypeMask;
        uint8_t* dst = p.image.data() + boff[i];
        if (t == kCOO) {
            const size_t k = p.coo_idx[o].size();
            const size_t ib = k * (p.compact ? 2 : 4);
            for (size_t q = 0; q < k; ++q) {
                if (p.compact) { const uint16_t w = (uint16_t)p.coo_idx[o][q]; std::memcpy(dst + 2 * q, &w, 2); }
                else std::memcpy(dst + 4 * q, &p.coo_idx[o][q], 4);
            }
            if (!unit) std::memcpy(dst + align_up(ib, 8), p.coo_val[o].data(), k * 8);
        } else if (t == kCSR) {
            const size_t k = p.csr_col[o].size();
            auto put_orig = [&](size_t pos, int v) {
                if (ob == 1) dst[pos] = (uint8_t)v;
                else { const uint16_t w = (uint16_t)v; std::memcpy(dst + pos * 2, &w, 2); }
            };
            for (int r = 0; r <= bs; ++r) put_orig(r, p.csr_rp[o][r]);
            for (size_t q = 0; q < k; ++q) put_orig(bs + 1 + q, p.csr_col[o][q]);
            if (!unit) std::memcpy(dst + align_up((bs + 1 + k) * ob, 8), p.csr_val[o].data(), k * 8);
        } else {
            if (unit) {
                const size_t mb = dense_mask_bytes(bs);
                for (int r = 0; r < bs; ++r) {
                    uint64_t mask = 0;
                    for (int c = 0; c < bs; ++c) if (p.dense_val[o][r * bs + c] != 0.0) mask |= 1ull << c;
                    std::memcpy(dst + r * mb, &mask, mb);
                }
            } else {
                std::memcpy(dst, p.dense_val[o].data(), (size_t)bs * bs * 8);
            }
        }
    }
    // x region left zero: the image is x-independent.
}

template <typename T>
long copy_out(const std::vector<T>& v, void* dst) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(T));
    return (long)v.size();
}

}  // namespace

extern "C" {

struct oracle_opts { int bs, waves, row_owned, force_coo64, dense_sixteenths, pattern_wave; };

void* oracle_build(int m, int n, int nnz, const int* row, const int* col, const double* val,
                   const oracle_opts* o, int comp) {
    Opts op{o->bs, o->waves, o->row_owned, o->force_coo64, o->dense_sixteenths, o->pattern_wave};
    return comp ? build_comp(m, n, nnz, row, col, val, op) : build_gather(m, n, nnz, row, col, val, op);
}

void oracle_free(void* h) { delete static_cast<Plan*>(h); }

double oracle_scalar(void* h, const char* name) {
    const Plan& p = *static_cast<Plan*>(h);
    const std::string s(name);
    if (s == "gb_m") return p.gb_m;
    if (s == "gb_n") return p.gb_n;
    if (s == "gb_nnzb") return p.nnzb;
    if (s == "gather_flag") return p.gather;
    if (s == "direct_coo_columns") return p.direct;
    if (s == "compact_coo_indices") return p.compact;
    if (s == "super_sparse_blocks") return p.super_sparse;
    if (s == "analyzed_blocks") return p.analyzed;
    if (s == "super_sparse_ratio") return p.super_ratio;
    if (s == "load_balanced") return p.balanced;
    if (s == "load_stddev_before") return p.sd_before;
    if (s == "load_stddev_after") return p.sd_after;
    return NAN;
}

// Copies the named array into dst (if non-null) and returns its element count.
// Payload streams, in final block order with pad blocks skipped:
//   "pl_idx": COO index words (u32) | CSR row_ptr[bs+1] then col[k] (i32)
//   "pl_val": COO values | CSR values | DENSE values[bs*bs] (f64)
long oracle_array(void* h, const char* name, void* dst) {
    Plan& p = *static_cast<Plan*>(h);
    const std::string s(name);
    if (s == "gb_row_idx") return copy_out(p.brow, dst);
    if (s == "gb_col_idx") return copy_out(p.bcol, dst);
    if (s == "gb_row_ptr") return copy_out(p.row_ptr, dst);
    if (s == "nnz_per_blk") return copy_out(p.nnzpb, dst);
    if (s == "type_per_blk") return copy_out(p.type, dst);
    if (s == "segment_rows") return copy_out(p.seg_rows, dst);
    if (s == "segment_begin") return copy_out(p.seg_begin, dst);
    if (s == "segment_end") return copy_out(p.seg_end, dst);
    if (s == "segment_direct") return copy_out(p.seg_direct, dst);
    if (s == "segment_uniform") return copy_out(p.seg_uniform, dst);
    if (s == "block_row_uniform") return copy_out(p.row_uniform, dst);
    if (s == "flat_row_ptr") return copy_out(p.flat_row_ptr, dst);
    if (s == "flat_col_idx") return copy_out(p.flat_col, dst);
    if (s == "pattern_row_ptr") return copy_out(p.pat_row_ptr, dst);
    if (s == "pattern_col_idx") return copy_out(p.pat_col, dst);
    if (s == "pattern_output_ptr") return copy_out(p.pat_out_ptr, dst);
    if (s == "pattern_output_rows") return copy_out(p.pat_out_rows, dst);
    if (s == "block_cols") return copy_out(p.block_cols, dst);
    if (s == "block_cols_offset") return copy_out(p.block_cols_off, dst);
    if (s == "bia_nnz") return copy_out(p.bia_nnz, dst);
    if (s == "bia_origin") return copy_out(p.bia_origin, dst);
    if (s == "bia_end") return copy_out(p.bia_end, dst);
    if (s == "pl_idx" || s == "pl_val") {
        std::vector<int32_t> idx;
        std::vector<double> val;
        for (int i = 0; i < p.nnzb; ++i) {
            if (p.nnzpb[i] <= 0) continue;
            const int o = p.bia_origin[i];
            const int t = p.type[i] & kTypeMask;
            if (t == kCOO) {
                for (uint32_t w : p.coo_idx[o]) idx.push_back((int32_t)w);
                val.insert(val.end(), p.coo_val[o].begin(), p.coo_val[o].end());
            } else if (t == kCSR) {
                idx.insert(idx.end(), p.csr_rp[o].begin(), p.csr_rp[o].end());
                idx.insert(idx.end(), p.csr_col[o].begin(), p.csr_col[o].end());
                val.insert(val.end(), p.csr_val[o].begin(), p.csr_val[o].end());
            } else {
                val.insert(val.end(), p.dense_val[o].begin(), p.dense_val[o].end());
            }
        }
        return s == "pl_idx" ? copy_out(idx, dst) : copy_out(val, dst);
    }
    if (s == "image") {
        if (p.image.empty()) pack(p);
        return copy_out(p.image, dst);
    }
    return -1;
}

// Byte offset of a slab region ("total" = slab size); -1 if absent.
long oracle_layout(void* h, const char* name) {
    Plan& p = *static_cast<Plan*>(h);
    if (p.image.empty()) pack(p);
    auto it = p.off.find(name);
    return it == p.off.end() ? -1 : (long)it->second;
}

// Dispatch policy (cb-spmv-dcu.hpp:1428-1450) as plain integers:
// [unit_values, use_pattern, use_unrolled, use_small_direct, launch_waves,
//  num_blocks(grid), aggregate_coo, pattern_count]
void oracle_policy(void* h, long* out) {
    Plan& p = *static_cast<Plan*>(h);
    if (p.image.empty()) pack(p);
    const long P = p.pat_row_ptr.empty() ? 0 : (long)p.pat_row_ptr.size() - 1;
    const bool pat = p.o.pattern_wave && p.unit_values && P > 0 && 2 * P < p.m;
    const bool unr = pat && (long)p.pat_col.size() >= P * 128;
    const bool small = p.o.row_owned && !pat && p.unit_values && p.gather && p.compact && p.nnz < 10000;
    const long waves = (pat && P < 1000) ? 9 : p.o.waves;
    const long grid = pat ? (P + waves - 1) / waves
                          : p.o.row_owned ? (long)p.seg_rows.size() : (p.nnzb + p.o.waves - 1) / p.o.waves;
    out[0] = p.unit_values; out[1] = pat; out[2] = unr; out[3] = small;
    out[4] = waves; out[5] = grid; out[6] = p.nnz >= 10000; out[7] = P;
}

// y = A x by the reference's sequential loop (utils.h:98-111): entries in
// input order, y[row] += val * x[col], FP64.  y is zeroed first.
void oracle_spmv_seq(int nnz, const int* row, const int* col, const double* val,
                     int m, const double* x, double* y) {
    for (int r = 0; r < m; ++r) y[r] = 0.0;
    for (int i = 0; i < nnz; ++i) y[row[i]] += val[i] * x[col[i]];
}

// Row-parallel CSR variant of the same sum (row-major input: entries of a row
// are contiguous, so every y[r] accumulates in the same order as above).
void oracle_spmv_csr_omp(int m, const long* row_ptr, const int* col, const double* val,
                         const double* x, double* y, int threads) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 256)
    for (int r = 0; r < m; ++r) {
        double s = 0.0;
        for (long k = row_ptr[r]; k < row_ptr[r + 1]; ++k) s += val[k] * x[col[k]];
        y[r] = s;
    }
}

// Median seconds per call of oracle_spmv_csr_omp over `reps` calls after one
// warm-up call; the caller builds the CSR arrays once, so only the SpMV
// loop is timed.
double oracle_time_csr_omp(int m, const long* row_ptr, const int* col, const double* val,
                           const double* x, double* y, int threads, int reps) {
    oracle_spmv_csr_omp(m, row_ptr, col, val, x, y, threads);
    std::vector<double> t;
    for (int k = 0; k < reps; ++k) {
        const auto a = std::chrono::steady_clock::now();
        oracle_spmv_csr_omp(m, row_ptr, col, val, x, y, threads);
        const auto b = std::chrono::steady_clock::now();
        t.push_back(std::chrono::duration<double>(b - a).count());
    }
    std::sort(t.begin(), t.end());
    return t.empty() ? 0.0 : t[t.size() / 2];
}

// The host's thread counts: omp_get_max_threads(),
// std::thread::hardware_concurrency() and the size of this process's CPU
// affinity mask.
void oracle_host_threads(int* omp_max, int* hw, int* affinity) {
    *omp_max = omp_get_max_threads();
    *hw = (int)std::thread::hardware_concurrency();
    cpu_set_t set;
    CPU_ZERO(&set);
    *affinity = sched_getaffinity(0, sizeof(set), &set) == 0 ? CPU_COUNT(&set) : -1;
}

}  // extern "C"
