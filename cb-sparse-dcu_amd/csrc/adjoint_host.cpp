// Host half of the adjoint: the transposed plan and the entry maps that tie
// input COO entries to stream entries.
//
// There is no reference counterpart (the reference computes y = A x only).
// A^T is built once, with the same builder and options, from the input COO
// reordered column-major (a stable counting sort on the column, so entries of
// one column keep their input order); its stream schedule then runs
// dx = A^T g with the forward kernel, and its SDDMM variant writes
// dvals = g_i x_j in the transposed plan's stream order ("stored order").
// grad_slot[k] is the stored-order index of input entry k.
#include <algorithm>
#include <new>

#include "cb_plan.hpp"
#include "cb_runtime.hpp"

namespace cbs {

// Per stream entry that holds input entries: its value's byte offset in the
// stream slab and its input entries (CSR over groups, input order inside).
int value_groups(const Plan& p, const std::vector<uint32_t>& slot, std::vector<uint32_t>& pos,
                 std::vector<uint32_t>& ptr, std::vector<uint32_t>& src) {
    if (p.s_slab.size() >= (1ull << 32)) return fail(CBSPMV_ERR_UNSUPPORTED, "stream slab above 4 GiB");
    const size_t nnz = slot.size();
    const uint32_t padded = p.items.empty() ? 0u : p.items.back().e_end;
    std::vector<uint32_t> cnt((size_t)padded + 1, 0);
    for (uint32_t s : slot) cnt[s + 1]++;
    for (uint32_t s = 0; s < padded; ++s) cnt[s + 1] += cnt[s];
    src.assign(nnz, 0);
    {
        std::vector<uint32_t> at(cnt.begin(), cnt.end() - 1);
        for (size_t k = 0; k < nnz; ++k) src[at[slot[k]]++] = (uint32_t)k;
    }
    pos.clear();
    ptr.clear();
    size_t item = 0;
    for (uint32_t s = 0; s < padded; ++s) {
        if (cnt[s + 1] == cnt[s]) continue;
        while (p.items[item].e_end <= s) ++item;
        pos.push_back((uint32_t)stream_value_offset(p, (uint32_t)item, s));
        ptr.push_back(cnt[s]);
    }
    ptr.push_back((uint32_t)nnz);
    return CBSPMV_OK;
}

}  // namespace cbs

using cbs::fail;

extern "C" {

int cbspmv_build_transpose(cbspmv_plan* plan) {
    if (!plan) return fail(CBSPMV_ERR_ARG, "null plan");
    if (plan->transpose) return CBSPMV_OK;
    const cbs::Plan& h = plan->host;
    const int m = h.m, n = h.n, nnz = h.nnz;
    if ((long long)m > (1ll << (30 - cbs::coord_bits(h.o.bs))))
        return fail(CBSPMV_ERR_UNSUPPORTED, "m above 2^(30 - log2 block_size): A^T cannot be built");
    try {
        // column-major order of the input entries, stable within a column
        std::vector<int> cptr((size_t)n + 1, 0);
        for (int c : h.in_col) cptr[c + 1]++;
        for (int c = 0; c < n; ++c) cptr[c + 1] += cptr[c];
        std::vector<int> perm((size_t)nnz);
        {
            std::vector<int> at(cptr.begin(), cptr.end() - 1);
            for (int k = 0; k < nnz; ++k) perm[at[h.in_col[k]]++] = k;
        }
        std::vector<int> trow((size_t)nnz), tcol((size_t)nnz);
        std::vector<double> tval((size_t)nnz);
        for (int q = 0; q < nnz; ++q) {
            trow[q] = h.in_col[perm[q]];
            tcol[q] = h.in_row[perm[q]];
            tval[q] = h.in_val[perm[q]];
        }
        cbspmv_plan* t = nullptr;
        // the backward launch (with the SDDMM stores) is fastest at 8 waves
        // per tile for every value type
        cbspmv_opts to = plan->opts;
        if (!to.stream_tile_waves) to.stream_tile_waves = cbs::kDefaultTileItemsOpt;
        int rc = cbs::build_plan_object(to, n, m, nnz, trow.data(), tcol.data(), tval.data(), false, &t);
        if (rc) return rc;
        std::vector<uint32_t> own;
        if (!cbs::stream_entry_slots(t->host, own) || !cbs::stream_entry_slots(plan->host, plan->slot_of)) {
            cbspmv_destroy(t);
            return fail(CBSPMV_ERR_STATE, "stream entries do not match the input entries");
        }
        plan->grad_slot.assign((size_t)nnz, 0);
        for (int q = 0; q < nnz; ++q) plan->grad_slot[perm[q]] = own[q];
        plan->transpose = t;
    } catch (const std::bad_alloc&) {
        return fail(CBSPMV_ERR_NOMEM, "out of host memory while building A^T");
    }
    return CBSPMV_OK;
}

long long cbspmv_stored_entries(const cbspmv_plan* plan) {
    if (!plan || !plan->transpose) return -1;
    const cbs::Plan& t = plan->transpose->host;
    return t.items.empty() ? 0 : (long long)t.items.back().e_end;
}

}  // extern "C"
