// Device half of the adjoint: dx = A^T g, dvals = g_i x_j (SDDMM over the
// stored pattern), and value replacement.  No reference counterpart.
//
// Backward modes (cbspmv_opts.backward_mode):
//   fused   one stream pass over the transposed plan writes dx and dvals in
//           stored order (the SDDMM reuses the g_i the product gathers); a
//           permutation pass then brings dvals to input order when asked;
//   direct  dx from the transposed plan's stream pass, dvals in input order
//           from a separate pass over the input coordinates.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "cb_runtime.hpp"
#include "kernels.hpp"

namespace cbs {
int value_groups(const Plan& p, const std::vector<uint32_t>& slot, std::vector<uint32_t>& pos,
                 std::vector<uint32_t>& ptr, std::vector<uint32_t>& src);

namespace dev {

__device__ __forceinline__ uint16_t to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40);
    return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// One thread per stream entry holding input entries: sum them (input order,
// FP64) and store the value in the entry's slot of the stream slab.
template <typename A, int DT>
__global__ void set_values(const A* vals, const uint32_t* pos, const uint32_t* ptr, const uint32_t* src,
                           uint32_t groups, uint8_t* slab) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < groups; g += gridDim.x * blockDim.x) {
        double s = 0.0;
        for (uint32_t q = ptr[g]; q < ptr[g + 1]; ++q) s += (double)vals[src[q]];
        uint8_t* dst = slab + pos[g];
        if constexpr (DT == CBSPMV_F64) *reinterpret_cast<double*>(dst) = s;
        else if constexpr (DT == CBSPMV_F32) *reinterpret_cast<float*>(dst) = (float)s;
        else *reinterpret_cast<uint16_t*>(dst) = to_bf16((float)s);
    }
}

template <typename A>
__global__ void gather_slots(const A* src, const uint32_t* idx, A* dst, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        dst[k] = src[idx[k]];
}

// dvals[k] = g[row[k]] * x[col[k]] over the input entries
template <typename A>
__global__ void sddmm_direct(const int* row, const int* col, const A* g, const A* x, A* dvals, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
        dvals[k] = g[row[k]] * x[col[k]];
}

}  // namespace dev

// Scratch of cbspmv_backward_stored_noalloc: the dx output the fused pass
// writes when the caller passes no dx buffer.  A file-scope device array,
// so the loader places one copy on every GPU the library runs on and keeps
// it until the process ends; no allocation call is made for it.
constexpr size_t kNoAllocScratchBytes = size_t(16) << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

// Device address of `scratch` on the current device, looked up once per
// device with that device current.  The host-side name `scratch` is a
// stand-in, not a device address: only the looked-up pointer may be passed
// to kernels.
std::mutex g_scratch_mu;
std::vector<void*> g_scratch_ptr;

int scratch_on_current_device(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)dev >= g_scratch_ptr.size()) g_scratch_ptr.resize((size_t)dev + 1, nullptr);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        if (!ptr) return fail(CBSPMV_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return CBSPMV_OK;
}

int grid_for(size_t n) { return (int)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 8192)); }

cbspmv_plan* ready_transpose(cbspmv_plan* plan, hipStream_t s, int& rc) {
    rc = CBSPMV_OK;
    if (!plan->transpose) {
        rc = fail(CBSPMV_ERR_STATE, "no transposed plan (call cbspmv_build_transpose first)");
        return nullptr;
    }
    if (!plan->dev.uploaded) {
        rc = fail(CBSPMV_ERR_STATE, "plan not uploaded (call cbspmv_upload first)");
        return nullptr;
    }
    if ((rc = upload_plan(*plan->transpose, s))) return nullptr;
    return plan->transpose;
}

// slot: the stream index in p of every input entry of the forward matrix
int ensure_groups(cbspmv_plan& p, const std::vector<uint32_t>& slot, hipStream_t s) {
    DevicePlan& d = p.dev;
    if (d.sv_pos.ptr || p.host.nnz == 0) return CBSPMV_OK;
    std::vector<uint32_t> pos, ptr, src;
    int rc = value_groups(p.host, slot, pos, ptr, src);
    if (rc) return rc;
    if ((rc = upload_vec(d.sv_pos, pos, s)) || (rc = upload_vec(d.sv_ptr, ptr, s)) || (rc = upload_vec(d.sv_src, src, s)))
        return rc;
    d.sv_groups = (uint32_t)pos.size();
    return CBSPMV_OK;
}

// Stored-order groups: values come as one per stored entry (the transposed
// plan's padded stream order).  Every stored entry that holds input entries
// is owned by the stream entry of p's plan that holds its first input entry
// (lowest input index); a group sums the stored entries it owns.  For the
// transposed plan itself that is the identity (each group reads its own
// stored entry).
int ensure_stored_groups(cbspmv_plan& p, const std::vector<uint32_t>& slot, const std::vector<uint32_t>& grad_slot,
                         bool identity, hipStream_t s) {
    DevicePlan& d = p.dev;
    if (d.ss_pos.ptr || p.host.nnz == 0) return CBSPMV_OK;
    std::vector<uint32_t> pos, ptr, src;
    int rc = value_groups(p.host, slot, pos, ptr, src);
    if (rc) return rc;
    std::vector<uint32_t> sptr(1, 0), ssrc;
    ssrc.reserve(pos.size());
    if (identity) {
        // slot == grad_slot here: a group's input entries share one stored entry
        for (size_t g = 0; g < pos.size(); ++g) {
            ssrc.push_back(grad_slot[src[ptr[g]]]);
            sptr.push_back((uint32_t)ssrc.size());
        }
    } else {
        size_t stored = 0;
        for (uint32_t e : grad_slot) stored = std::max<size_t>(stored, (size_t)e + 1);
        std::vector<uint32_t> owner(stored, 0xffffffffu);
        for (size_t k = grad_slot.size(); k-- > 0;) owner[grad_slot[k]] = (uint32_t)k;   // lowest k wins
        for (size_t g = 0; g < pos.size(); ++g) {
            for (uint32_t q = ptr[g]; q < ptr[g + 1]; ++q) {
                const uint32_t k = src[q];
                if (owner[grad_slot[k]] == k) ssrc.push_back(grad_slot[k]);
            }
            sptr.push_back((uint32_t)ssrc.size());
        }
    }
    if ((rc = upload_vec(d.ss_pos, pos, s)) || (rc = upload_vec(d.ss_ptr, sptr, s)) || (rc = upload_vec(d.ss_src, ssrc, s)))
        return rc;
    d.ss_groups = (uint32_t)pos.size();
    return CBSPMV_OK;
}

int launch_group_values(cbspmv_plan& p, const void* vals, const DeviceBuf& gpos, const DeviceBuf& gptr,
                        const DeviceBuf& gsrc, uint32_t groups, hipStream_t s) {
    DevicePlan& d = p.dev;
    if (groups == 0) return CBSPMV_OK;
    const int grid = grid_for(groups);
    const auto* pos = static_cast<const uint32_t*>(gpos.ptr);
    const auto* ptr = static_cast<const uint32_t*>(gptr.ptr);
    const auto* src = static_cast<const uint32_t*>(gsrc.ptr);
    auto* slab = static_cast<uint8_t*>(d.key.ptr);
    switch (p.host.o.dtype) {
        case CBSPMV_F64:
            hipLaunchKernelGGL((dev::set_values<double, CBSPMV_F64>), dim3(grid), dim3(256), 0, s,
                               static_cast<const double*>(vals), pos, ptr, src, groups, slab);
            break;
        case CBSPMV_F32:
            hipLaunchKernelGGL((dev::set_values<float, CBSPMV_F32>), dim3(grid), dim3(256), 0, s,
                               static_cast<const float*>(vals), pos, ptr, src, groups, slab);
            break;
        default:
            hipLaunchKernelGGL((dev::set_values<float, CBSPMV_BF16>), dim3(grid), dim3(256), 0, s,
                               static_cast<const float*>(vals), pos, ptr, src, groups, slab);
    }
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? CBSPMV_OK : hip_fail(e, "set_values launch");
}

int launch_set_values(cbspmv_plan& p, const void* vals, hipStream_t s) {
    const DevicePlan& d = p.dev;
    return launch_group_values(p, vals, d.sv_pos, d.sv_ptr, d.sv_src, d.sv_groups, s);
}

// A device buffer of at least `bytes` owned by the plan (grown on demand).
int plan_scratch(DeviceBuf& b, size_t bytes, hipStream_t s, void** out) {
    if (b.bytes < bytes) {
        const int rc = upload_bytes(b, nullptr, bytes, s);
        if (rc) return rc;
    }
    *out = b.ptr;
    return CBSPMV_OK;
}

}  // namespace
}  // namespace cbs

using cbs::fail;

extern "C" {

int cbspmv_spmv_transpose(cbspmv_plan* plan, const void* d_g, void* d_dx, void* stream) {
    if (!plan) return fail(CBSPMV_ERR_ARG, "null plan");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    cbspmv_plan* t = cbs::ready_transpose(plan, s, rc);
    if (!t) return rc;
    if ((t->host.m > 0 && !d_dx) || (t->host.nnz > 0 && !d_g)) return fail(CBSPMV_ERR_ARG, "null vector");
    return cbs::run_spmv(*t, d_g, d_dx, s);
}

int cbspmv_backward_stored(cbspmv_plan* plan, const void* d_g, const void* d_x, void* d_dx,
                           void* d_dvals_stored, void* stream) {
    if (!plan) return fail(CBSPMV_ERR_ARG, "null plan");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    cbspmv_plan* t = cbs::ready_transpose(plan, s, rc);
    if (!t) return rc;
    const cbs::Plan& th = t->host;
    if (th.nnz > 0 && (!d_g || !d_x || !d_dvals_stored)) return fail(CBSPMV_ERR_ARG, "null vector");
    if (reinterpret_cast<uintptr_t>(d_dvals_stored) % 16) return fail(CBSPMV_ERR_ARG, "dvals must be 16-byte aligned");
    if (th.nnz == 0 || th.m == 0) {
        if (d_dx && th.m) return cbs::run_spmv(*t, d_g, d_dx, s);
        return CBSPMV_OK;
    }
    void* dx = d_dx;
    if (!dx && (rc = cbs::plan_scratch(plan->dev.dvals_stored, (size_t)th.m * cbs::vec_elem_bytes(th), s, &dx)))
        return rc;
    return cbs::run_stream(*t, d_g, dx, s, d_x, d_dvals_stored);
}

int cbspmv_backward_stored_noalloc(cbspmv_plan* plan, const void* d_g, const void* d_x, void* d_dx,
                                   void* d_dvals_stored, void* stream) {
    if (!plan) return fail(CBSPMV_ERR_ARG, "null plan");
    cbspmv_plan* t = plan->transpose;
    if (!t) return fail(CBSPMV_ERR_STATE, "no transposed plan (call cbspmv_build_transpose first)");
    const cbs::Plan& th = t->host;
    if (th.nnz > 0 && (!d_g || !d_x || !d_dvals_stored)) return fail(CBSPMV_ERR_ARG, "null vector");
    if (reinterpret_cast<uintptr_t>(d_dvals_stored) % 16) return fail(CBSPMV_ERR_ARG, "dvals must be 16-byte aligned");
    const size_t dx_bytes = (size_t)th.m * cbs::vec_elem_bytes(th);
    if (!d_dx && dx_bytes > cbs::kNoAllocScratchBytes)
        return fail(CBSPMV_ERR_ARG, "n times the vector element size is above the 16 MiB scratch: pass a dx buffer");
    int dev = -1;
    if (!plan->dev.uploaded || !t->dev.uploaded || hipGetDevice(&dev) != hipSuccess || dev != plan->dev.device ||
        dev != t->dev.device)
        return fail(CBSPMV_ERR_STATE, "the plan and its transpose must be uploaded to the current device "
                                      "(cbspmv_upload after cbspmv_build_transpose)");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    if (th.nnz == 0 || th.m == 0) return d_dx && th.m ? cbs::run_spmv(*t, d_g, d_dx, s) : CBSPMV_OK;
    void* dx = d_dx;
    int rc;
    if (!dx && (rc = cbs::scratch_on_current_device(&dx))) return rc;
    return cbs::run_stream(*t, d_g, dx, s, d_x, d_dvals_stored);
}

int cbspmv_backward(cbspmv_plan* plan, const void* d_g, const void* d_x, void* d_dx, void* d_dvals,
                    void* stream) {
    if (!plan) return fail(CBSPMV_ERR_ARG, "null plan");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    cbspmv_plan* t = cbs::ready_transpose(plan, s, rc);
    if (!t) return rc;
    const cbs::Plan& h = plan->host;
    const cbs::Plan& th = t->host;
    if (!d_dvals) {
        if (!d_dx) return CBSPMV_OK;
        return cbs::run_spmv(*t, d_g, d_dx, s);
    }
    if (h.nnz > 0 && (!d_g || !d_x)) return fail(CBSPMV_ERR_ARG, "null vector");
    if (h.nnz == 0) return d_dx ? cbs::run_spmv(*t, d_g, d_dx, s) : CBSPMV_OK;
    const bool f64 = h.o.dtype == CBSPMV_F64;
    const size_t eb = cbs::vec_elem_bytes(h);
    const int mode = plan->opts.backward_mode;
    cbs::DevicePlan& d = plan->dev;
    if (mode == CBSPMV_BACKWARD_DIRECT) {
        if (!d.in_row.ptr) {
            if ((rc = cbs::upload_vec(d.in_row, h.in_row, s)) || (rc = cbs::upload_vec(d.in_col, h.in_col, s))) return rc;
        }
        if (d_dx && (rc = cbs::run_spmv(*t, d_g, d_dx, s))) return rc;
        const int grid = cbs::grid_for((size_t)h.nnz);
        const int* r = static_cast<const int*>(d.in_row.ptr);
        const int* c = static_cast<const int*>(d.in_col.ptr);
        if (f64)
            hipLaunchKernelGGL(cbs::dev::sddmm_direct<double>, dim3(grid), dim3(256), 0, s, r, c,
                               static_cast<const double*>(d_g), static_cast<const double*>(d_x),
                               static_cast<double*>(d_dvals), (uint32_t)h.nnz);
        else
            hipLaunchKernelGGL(cbs::dev::sddmm_direct<float>, dim3(grid), dim3(256), 0, s, r, c,
                               static_cast<const float*>(d_g), static_cast<const float*>(d_x),
                               static_cast<float*>(d_dvals), (uint32_t)h.nnz);
        const hipError_t e = hipGetLastError();
        return e == hipSuccess ? CBSPMV_OK : cbs::hip_fail(e, "sddmm launch");
    }
    // fused: stored-order dvals into the plan's buffer, then to input order
    const size_t stored = th.items.empty() ? 0 : th.items.back().e_end;
    if (!d.grad_slot.ptr && (rc = cbs::upload_vec(d.grad_slot, plan->grad_slot, s))) return rc;
    const size_t dvb = (stored * eb + 255) / 256 * 256;
    void* raw = nullptr;
    if ((rc = cbs::plan_scratch(d.dvals_stored, dvb + (d_dx ? 0 : (size_t)th.m * eb), s, &raw))) return rc;
    uint8_t* buf = static_cast<uint8_t*>(raw);
    void* dx = d_dx ? d_dx : buf + dvb;
    if ((rc = cbs::run_stream(*t, d_g, dx, s, d_x, buf))) return rc;
    const int grid = cbs::grid_for((size_t)h.nnz);
    const uint32_t* idx = static_cast<const uint32_t*>(d.grad_slot.ptr);
    if (f64)
        hipLaunchKernelGGL(cbs::dev::gather_slots<double>, dim3(grid), dim3(256), 0, s,
                           reinterpret_cast<const double*>(buf), idx, static_cast<double*>(d_dvals), (uint32_t)h.nnz);
    else
        hipLaunchKernelGGL(cbs::dev::gather_slots<float>, dim3(grid), dim3(256), 0, s,
                           reinterpret_cast<const float*>(buf), idx, static_cast<float*>(d_dvals), (uint32_t)h.nnz);
    const hipError_t e = hipGetLastError();
    return e == hipSuccess ? CBSPMV_OK : cbs::hip_fail(e, "gather launch");
}

int cbspmv_set_values_stored(cbspmv_plan* plan, const void* d_vals_stored, void* stream) {
    if (!plan) return fail(CBSPMV_ERR_ARG, "null plan");
    if (!plan->transpose) return fail(CBSPMV_ERR_STATE, "no transposed plan (call cbspmv_build_transpose first)");
    if (!plan->dev.uploaded) return fail(CBSPMV_ERR_STATE, "plan not uploaded (call cbspmv_upload first)");
    if (plan->host.nnz == 0) return CBSPMV_OK;
    if (!d_vals_stored) return fail(CBSPMV_ERR_ARG, "null values");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    cbspmv_plan* t = cbs::ready_transpose(plan, s, rc);
    if (!t) return rc;
    if ((rc = cbs::ensure_stored_groups(*plan, plan->slot_of, plan->grad_slot, false, s)) ||
        (rc = cbs::ensure_stored_groups(*t, plan->grad_slot, plan->grad_slot, true, s)))
        return rc;
    const cbs::DevicePlan& d = plan->dev;
    const cbs::DevicePlan& td = t->dev;
    if ((rc = cbs::launch_group_values(*plan, d_vals_stored, d.ss_pos, d.ss_ptr, d.ss_src, d.ss_groups, s)) ||
        (rc = cbs::launch_group_values(*t, d_vals_stored, td.ss_pos, td.ss_ptr, td.ss_src, td.ss_groups, s)))
        return rc;
    plan->values_replaced = true;
    t->values_replaced = true;
    cbs::drop_split(*plan);   // the full stream holds the new values
    return CBSPMV_OK;
}

int cbspmv_set_values(cbspmv_plan* plan, const void* d_vals, void* stream) {
    if (!plan) return fail(CBSPMV_ERR_ARG, "null plan");
    if (!plan->dev.uploaded) return fail(CBSPMV_ERR_STATE, "plan not uploaded (call cbspmv_upload first)");
    if (plan->host.nnz == 0) return CBSPMV_OK;
    if (!d_vals) return fail(CBSPMV_ERR_ARG, "null values");
    const hipStream_t s = static_cast<hipStream_t>(stream);
    int rc;
    if (plan->slot_of.empty() && !cbs::stream_entry_slots(plan->host, plan->slot_of))
        return fail(CBSPMV_ERR_STATE, "stream entries do not match the input entries");
    if ((rc = cbs::ensure_groups(*plan, plan->slot_of, s)) || (rc = cbs::launch_set_values(*plan, d_vals, s)))
        return rc;
    plan->values_replaced = true;
    cbs::drop_split(*plan);   // the full stream holds the new values
    if (plan->transpose) {
        // A^T's stream index of input entry k is grad_slot[k], so its groups
        // read the same input-order values
        cbspmv_plan* t = cbs::ready_transpose(plan, s, rc);
        if (!t) return rc;
        if ((rc = cbs::ensure_groups(*t, plan->grad_slot, s)) || (rc = cbs::launch_set_values(*t, d_vals, s))) return rc;
        t->values_replaced = true;
    }
    return CBSPMV_OK;
}

}  // extern "C"
