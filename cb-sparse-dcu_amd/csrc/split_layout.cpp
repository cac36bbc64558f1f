// Scattered-entry split, host layout: which entries are scattered, and the
// order the sweep kernel reads them in (sweep.hip).  Host code only, no HIP:
// the split's plans and policy are in split.cpp.
//
// A line read kScatteredReads times or fewer in a group is a scattered
// line.  Uniformly scattered columns meet in a line by chance: on the
// 1,048,576^2 stencil / power-law matrix ~42% of them share their line with
// another one in their group, and almost none with three others; a line the
// stream stages (the diagonal band) is read hundreds of times.
#include <algorithm>
#include <utility>
#include <vector>

#include "cb_plan.hpp"

namespace cbs {
namespace {

// Entries of one sweep block at most (a row with more is a block alone).
constexpr int64_t kSweepBlockEntries = 8000;
}  // namespace

// Rows are cut into groups of about nnz / tiles entries at row boundaries
// (the stream schedule's tiles), and an entry is scattered when at most
// kScatteredReads entries of its group read its x line (line_elems
// consecutive columns).
std::vector<uint8_t> scattered_entries(int nnz, const int* row, const int* col, int tiles, uint32_t line_elems) {
    std::vector<uint8_t> far((size_t)nnz, 0);
    if (nnz == 0) return far;
    const int64_t target = std::max<int64_t>(1, (int64_t)nnz / std::max(1, tiles));
    std::vector<int64_t> cut{0};
    for (int64_t e = target; e < nnz; e += target) {
        while (e < nnz && row[e] == row[e - 1]) ++e;
        if (e >= nnz) break;
        cut.push_back(e);
    }
    cut.push_back(nnz);
    const long long groups = (long long)cut.size() - 1;
#pragma omp parallel
    {
        std::vector<std::pair<uint32_t, uint32_t>> ln;   // (line, entry)
#pragma omp for schedule(dynamic, 1)
        for (long long g = 0; g < groups; ++g) {
            ln.clear();
            for (int64_t e = cut[g]; e < cut[g + 1]; ++e) ln.push_back({(uint32_t)col[e] / line_elems, (uint32_t)e});
            std::sort(ln.begin(), ln.end());
            for (size_t q = 0; q < ln.size();) {
                size_t r = q + 1;
                while (r < ln.size() && ln[r].first == ln[q].first) ++r;
                if (r - q <= kScatteredReads)
                    for (size_t t = q; t < r; ++t) far[ln[t].second] = 1;
                q = r;
            }
        }
    }
    return far;
}

// The sweep layout of the scattered entries fe (input indices, row-major).
// Block and wave boundaries are found in one pass; the waves' entries are
// then ordered block by block in parallel, each into its own output range.
SweepPart sweep_layout(const std::vector<uint32_t>& fe, const int* row, const int* col, const double* val,
                       int col_bits, int max_rows) {
    SweepPart sw;
    sw.col_bits = col_bits;
    sw.max_rows = max_rows;
    const int64_t nf = (int64_t)fe.size();
    auto next_row = [&](int64_t k) {   // first entry after k's row
        const int r = row[fe[k]];
        while (k < nf && row[fe[k]] == r) ++k;
        return k;
    };
    std::vector<int64_t> cuts;   // per block: kSweepWaves + 1 boundaries into fe (shared between blocks)
    for (int64_t i = 0; i < nf;) {
        const int r0 = row[fe[i]];
        int64_t j = i;
        while (j < nf && row[fe[j]] - r0 < max_rows) {
            const int64_t k = next_row(j);
            if (j > i && k - i > kSweepBlockEntries) break;
            j = k;
        }
        // block: entries [i, j), rows [r0, row of entry j - 1]; its waves cut
        // at row boundaries near equal shares of its entries
        sw.blk_row.push_back(r0);
        sw.blk_span.push_back(row[fe[j - 1]] - r0 + 1);
        int64_t a = i;
        for (int w = 0; w < kSweepWaves; ++w) {
            cuts.push_back(a);
            int64_t b = j;
            if (w + 1 < kSweepWaves) {
                b = std::max(a, i + (j - i) * (w + 1) / kSweepWaves);
                if (b > i && b < j && row[fe[b]] == row[fe[b - 1]]) b = next_row(b);
                b = std::min(b, j);
            }
            a = b;
        }
        i = j;
    }
    cuts.push_back(nf);
    const long long waves = (long long)cuts.size() - 1;
    sw.wave_off.assign(cuts.begin(), cuts.end());
    sw.key.resize((size_t)nf);
    sw.val.resize((size_t)nf);
    sw.src.resize((size_t)nf);
#pragma omp parallel
    {
        std::vector<uint32_t> grp;
#pragma omp for schedule(dynamic, 16)
        for (long long w = 0; w < waves; ++w) {
            const int64_t a = cuts[w], b = cuts[w + 1];
            const int r0 = sw.blk_row[(size_t)(w / kSweepWaves)];
            grp.assign(fe.begin() + a, fe.begin() + b);
            std::stable_sort(grp.begin(), grp.end(), [&](uint32_t x, uint32_t y) {
                return col[x] != col[y] ? col[x] < col[y] : row[x] < row[y];
            });
            for (size_t q = 0; q < grp.size(); q += 64)
                std::stable_sort(grp.begin() + q, grp.begin() + std::min(grp.size(), q + 64),
                                 [&](uint32_t x, uint32_t y) { return row[x] < row[y]; });
            for (size_t t = 0; t < grp.size(); ++t) {
                const uint32_t e = grp[t];
                sw.key[(size_t)a + t] = ((uint32_t)(row[e] - r0) << col_bits) | (uint32_t)col[e];
                sw.val[(size_t)a + t] = val[e];
                sw.src[(size_t)a + t] = e;
            }
        }
    }
    return sw;
}

}  // namespace cbs
