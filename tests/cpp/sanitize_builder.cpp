// Host-only driver for the AddressSanitizer / UndefinedBehaviorSanitizer
// build of the CB builder, the packers and the scattered-entry split's host
// layout (builder.cpp, packer.cpp, split_layout.cpp), run by
// tests/test_sanitize_builder.py.  It feeds them malformed and edge-case COO
// streams (the reference builder has no error path for these and its
// behaviour on them is undefined, coo2block.h:77-121) and a seeded fuzz of
// small matrices over every option combination, checks the status codes,
// and exercises every export.  Any out-of-bounds access, use after free or
// undefined behaviour aborts the sanitizer build.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <string>
#include <vector>

#include "cb_plan.hpp"

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
    if (!ok) {
        std::printf("FAIL: %s\n", what.c_str());
        ++failures;
    }
}

struct Coo {
    int m = 0, n = 0;
    std::vector<int> row, col;
    std::vector<double> val;
};

const char* kExports[] = {
    "gb_row_idx", "gb_col_idx", "gb_row_ptr", "nnz_per_blk", "type_per_blk", "segment_rows", "segment_begin",
    "segment_end", "segment_direct", "segment_uniform", "block_row_uniform", "flat_row_ptr", "flat_col_idx",
    "pattern_row_ptr", "pattern_col_idx", "pattern_output_ptr", "pattern_output_rows", "block_cols",
    "block_cols_offset", "bia_nnz", "bia_origin", "bia_end", "pl_idx", "pl_val", "image", "stream_key",
    "stream_val", "stream_slab", "stream_items", "stream_tiles", "stream_tile_lines", "stream_empty_rows",
    "stream_value_pos", "stream_col_bits", "stream_panels", "stream_panel_items", "stream_panel_mask", "policy"};

struct Opt {
    int bs, row_owned, builder, dtype, panels, tile_items, tile_kib, force_coo64, item_nnz;
};

// Builds, packs and exports; returns the builder's status.
int run(const Coo& a, const Opt& o) {
    cbs::Plan p;
    p.o.bs = o.bs;
    p.o.row_owned = o.row_owned;
    p.o.builder = o.builder;
    p.o.dtype = o.dtype;
    p.o.panels = o.panels;
    p.o.tile_items = o.tile_items;
    p.o.tile_bytes = o.tile_kib * 1024;
    p.o.force_coo64 = o.force_coo64;
    p.o.item_nnz = o.item_nnz;
    std::string err;
    const int nnz = (int)a.row.size();
    const int rc = cbs::build_plan(p, a.m, a.n, nnz, a.row.data(), a.col.data(), a.val.data(), err);
    if (rc != CBSPMV_OK) return rc;
    cbs::pack_canonical(p);
    cbs::build_stream(p);
    for (const char* name : kExports) {
        int status = CBSPMV_OK;
        const long cnt = cbs::export_array(p, name, nullptr, 0, status);
        expect(status == CBSPMV_OK && cnt >= 0, std::string("export ") + name);
        if (cnt > 0) {
            std::vector<unsigned char> buf((size_t)cnt * 8 + 8);
            status = CBSPMV_OK;
            cbs::export_array(p, name, buf.data(), (size_t)cnt, status);
        }
    }
    std::vector<uint32_t> slot;
    expect(cbs::stream_entry_slots(p, slot), "stream entries hold the input entries");
    // the scattered-entry split's host layout of the same entries: every
    // scattered entry once, decoding back to its row and column
    int cb = 1;
    while ((1ll << cb) < (long long)a.n) ++cb;
    const int max_rows = cb >= 26 ? 0 : std::min(cbs::kSweepMaxRows, 1 << (32 - cb));
    for (int tiles : {1, 7, 256}) {
        const std::vector<uint8_t> far = cbs::scattered_entries(nnz, a.row.data(), a.col.data(), tiles, 16);
        std::vector<uint32_t> fe;
        for (int e = 0; e < nnz; ++e)
            if (far[e]) fe.push_back((uint32_t)e);
        if (fe.empty() || max_rows < 64) continue;
        const cbs::SweepPart sw = cbs::sweep_layout(fe, a.row.data(), a.col.data(), a.val.data(), cb, max_rows);
        const size_t nb = sw.blocks();
        expect(sw.key.size() == fe.size() && sw.src.size() == fe.size() && sw.val.size() == fe.size(), "split sizes");
        expect(sw.wave_off.size() == nb * cbs::kSweepWaves + 1 && sw.wave_off.back() == fe.size(), "split wave table");
        std::vector<uint32_t> seen(sw.src);
        std::sort(seen.begin(), seen.end());
        expect(seen == fe, "split holds each scattered entry once");
        for (size_t w = 0; w + 1 < sw.wave_off.size(); ++w) {
            const size_t bl = w / cbs::kSweepWaves;
            for (uint32_t k = sw.wave_off[w]; k < sw.wave_off[w + 1]; ++k) {
                const uint32_t e = sw.src[k];
                const int r = sw.blk_row[bl] + (int)(sw.key[k] >> cb);
                const int c = (int)(sw.key[k] & ((1u << cb) - 1u));
                if (r != a.row[e] || c != a.col[e] || (int)(sw.key[k] >> cb) >= sw.blk_span[bl] ||
                    sw.blk_span[bl] > max_rows) {
                    expect(false, "split entry decodes to its input entry");
                    return rc;
                }
            }
        }
    }
    return rc;
}

Opt base() { return Opt{16, 1, CBSPMV_BUILDER_GATHER, CBSPMV_F64, 1, 16, 96, 0, 0}; }

Coo random_coo(std::mt19937& g, int m, int n, int per_row, bool dup, bool unit) {
    Coo a;
    a.m = m;
    a.n = n;
    std::uniform_int_distribution<int> cd(0, std::max(0, n - 1));
    std::poisson_distribution<int> pd(per_row);
    std::uniform_real_distribution<double> vd(-1.0, 1.0);
    for (int r = 0; r < m && n > 0; ++r) {
        const int k = std::min(pd(g), dup ? 4 * n : n);
        std::vector<int> cols;
        for (int q = 0; q < k; ++q) cols.push_back(cd(g));
        if (!dup) {
            std::sort(cols.begin(), cols.end());
            cols.erase(std::unique(cols.begin(), cols.end()), cols.end());
            std::shuffle(cols.begin(), cols.end(), g);
        }
        for (int c : cols) {
            a.row.push_back(r);
            a.col.push_back(c);
            a.val.push_back(unit ? 1.0 : vd(g));
        }
    }
    return a;
}

}  // namespace

int main() {
    const Opt o = base();
    // ---- malformed streams: refused with a status, nothing read out of bounds
    {
        Coo a;
        a.m = 10; a.n = 7;
        a.row = {0, 3, 10}; a.col = {1, 2, 0}; a.val = {1, 2, 3};
        expect(run(a, o) == CBSPMV_ERR_RANGE, "row == m");
        a.row = {0, 3, 9}; a.col = {1, 7, 0};
        expect(run(a, o) == CBSPMV_ERR_RANGE, "col == n");
        a.row = {-1, 3, 9}; a.col = {1, 2, 0};
        expect(run(a, o) == CBSPMV_ERR_RANGE, "negative row");
        a.row = {0, 3, 9}; a.col = {1, -5, 0};
        expect(run(a, o) == CBSPMV_ERR_RANGE, "negative col");
        a.row = {0, 5, 3}; a.col = {1, 2, 0};
        expect(run(a, o) == CBSPMV_ERR_ORDER, "rows not non-decreasing");
    }
    // ---- edge shapes
    {
        Coo a;
        a.m = 5; a.n = 5;
        expect(run(a, o) == CBSPMV_OK, "nnz = 0");
        a.m = 0; a.n = 5;
        expect(run(a, o) == CBSPMV_OK, "m = 0");
        a.m = 5; a.n = 0;
        expect(run(a, o) == CBSPMV_OK, "n = 0");
        a.m = 37; a.n = 53;
        a.row = {36}; a.col = {52}; a.val = {2.5};
        for (int bs : {2, 4, 8, 16, 32, 64}) {
            Opt b = o;
            b.bs = bs;
            expect(run(a, b) == CBSPMV_OK, "one entry in the last partial block, bs " + std::to_string(bs));
        }
        // one (row, col) repeated many times
        a.row.assign(1000, 3); a.col.assign(1000, 4); a.val.assign(1000, 0.5);
        expect(run(a, o) == CBSPMV_OK, "1000 duplicates");
        // one row holding every column (cut into pieces), then empty rows
        a.m = 300; a.n = 40000;
        a.row.assign(40000, 150); a.col.resize(40000); a.val.assign(40000, 1.0);
        for (int c = 0; c < 40000; ++c) a.col[c] = c;
        Opt s = o; s.item_nnz = 256;
        expect(run(a, s) == CBSPMV_OK, "long row");
        // a full dense matrix with zero values inside (DENSE blocks)
        a.m = a.n = 64;
        a.row.clear(); a.col.clear(); a.val.clear();
        for (int r = 0; r < 64; ++r)
            for (int c = 0; c < 64; ++c) { a.row.push_back(r); a.col.push_back(c); a.val.push_back((r * c) % 3 ? 1.0 : 0.0); }
        for (int bs : {8, 16, 32, 64}) {
            Opt b = o; b.bs = bs; b.builder = CBSPMV_BUILDER_COMP;
            expect(run(a, b) == CBSPMV_OK, "dense, bs " + std::to_string(bs));
        }
    }
    // ---- switch points of the encodings: n = 4095/4096/65535/65536, nnz 9999/10000
    {
        std::mt19937 g(7);
        for (int n : {4095, 4096, 65535, 65536}) {
            Coo a = random_coo(g, 300, n, 20, false, false);
            expect(run(a, o) == CBSPMV_OK, "n = " + std::to_string(n));
            Coo u = random_coo(g, 300, n, 20, false, true);
            expect(run(u, o) == CBSPMV_OK, "unit, n = " + std::to_string(n));
        }
        for (int target : {9999, 10000}) {
            Coo a = random_coo(g, 2000, 3000, 5, false, true);
            a.row.resize(std::min<size_t>(a.row.size(), (size_t)target));
            a.col.resize(a.row.size());
            a.val.resize(a.row.size());
            expect(run(a, o) == CBSPMV_OK, "nnz = " + std::to_string(a.row.size()));
        }
    }
    // ---- seeded fuzz over the option space
    {
        std::mt19937 g(12345);
        const int bss[] = {2, 4, 8, 16, 32, 64};
        for (int t = 0; t < 240; ++t) {
            std::uniform_int_distribution<int> dm(0, 700), dn(1, 3000), dk(0, 30);
            const int m = dm(g), n = dn(g);
            Coo a = random_coo(g, m, n, dk(g), (t % 3) == 0, (t % 5) == 0);
            Opt b;
            b.bs = bss[t % 6];
            b.row_owned = (t / 6) % 2;
            b.builder = (t / 12) % 2 ? CBSPMV_BUILDER_COMP : CBSPMV_BUILDER_GATHER;
            b.dtype = (t / 24) % 3;
            b.panels = (t % 7) == 0 ? 2 << (t % 3) : 1;
            b.tile_items = b.panels > 1 ? 4 : 1 + (t % 16);
            b.tile_kib = 4 + (t * 37) % 150;
            b.force_coo64 = (t % 11) == 0;
            b.item_nnz = (t % 4) == 0 ? 64 * (1 + t % 8) : 0;
            expect(run(a, b) == CBSPMV_OK, "fuzz case " + std::to_string(t));
        }
    }
    std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
    return failures ? 1 : 0;
}
