// PyTorch-ROCm extension module xrecon_amd._C: the counterpart of the reference's
// pybind11 module xrecon_C (src/api/torch_binding.cpp:615-638).  It validates tensors,
// unpacks the options vector into host descriptors, allocates outputs and workspaces
// through torch's caching allocator, and calls the C-ABI of libxrecon_amd.so on the
// stream it is given.  It contains no compute of its own.
//
// Differences from the reference binding, each deliberate:
//  * geometry is read on the host once per plan (options may live on the CPU), instead
//    of 18 device->host .item() reads per call (torch_binding.cpp:85-108);
//  * work goes to the caller's stream (the reference uses the legacy default stream,
//    src/kernels/forward3d_flat.cu:506), so calls can be captured into a graph;
//  * CPU tensors run the library's host executor (the same per-thread bodies as the GPU
//    kernels on the host's cores), where the reference runs its std::thread CPU kernels
//    (torch_binding.cpp:421-425); GPU tensors always run the gfx950 kernels.
#include <torch/extension.h>
#include <c10/hip/HIPStream.h>

#include <cstring>
#include <list>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/xrecon_amd.h"

namespace {

[[noreturn]] void raise_xr(int rc, const char *what)
{
    TORCH_CHECK(false, what, " failed (code ", rc, "): ", xr_last_error());
}

std::vector<float> host_floats(const torch::Tensor &t, int64_t min_numel, const char *name)
{
    TORCH_CHECK(t.numel() >= min_numel, name, " tensor must have at least ", min_numel, " elements.");
    auto c = t.detach().to(torch::kCPU, torch::kFloat32).contiguous();
    return std::vector<float>(c.data_ptr<float>(), c.data_ptr<float>() + c.numel());
}

struct Desc {
    xr_volume3d vol{};
    xr_geometry3d geo{};
    bool is2d = false;
    int W = 0, H = 0, D = 0, Nu = 0, Nv = 0;
};

// 3D options layout: xrecon/projector/projector3d.py:98-107, unpacked as in
// src/api/torch_binding.cpp:82-124.
Desc unpack3d(const std::vector<float> &o)
{
    TORCH_CHECK(o.size() >= 18, "3D option tensor must have at least 18 elements.");
    Desc d;
    d.W = (int)o[0]; d.H = (int)o[1]; d.D = (int)o[2]; d.Nu = (int)o[3]; d.Nv = (int)o[4];
    const int type = (int)o[17];
    TORCH_CHECK(type == XR_PARALLEL || type == XR_FLAT || type == XR_ARC, "3D scan_type must be parallel, flat, or arc");
    d.vol = xr_volume3d{d.W, d.H, d.D, o[5], o[6], o[7], o[12], o[13], o[14]};
    d.geo = xr_geometry3d{type, d.Nu, d.Nv, o[8], o[9], o[15], o[16], o[10], o[11]};
    return d;
}

// 2D options layout: xrecon/projector/projector.py:106-110, unpacked as in
// src/api/torch_binding.cpp:62-80.
Desc unpack2d(const std::vector<float> &o)
{
    TORCH_CHECK(o.size() >= 12, "Option tensor must have at least 12 elements.");
    Desc d;
    d.is2d = true;
    d.W = (int)o[0]; d.H = (int)o[1]; d.D = 1; d.Nu = (int)o[2]; d.Nv = 1;
    const int type = (int)o[11];
    TORCH_CHECK(type == XR_PARALLEL || type == XR_FLAT || type == XR_ARC, "2D scan_type must be parallel, flat, or arc");
    xr_volume2d v2{d.W, d.H, o[3], o[4], o[8], o[9]};
    xr_geometry2d g2{type, d.Nu, o[5], o[10], o[6], o[7]};
    const int rc = xr_geometry2d_as_3d(&v2, &g2, &d.vol, &d.geo);
    if (rc) raise_xr(rc, "xr_geometry2d_as_3d");
    return d;
}

bool volume_in(int op)
{
    return op == XR_OP_PROJECTION || op == XR_OP_BACKPROJECTION_T || op == XR_OP_BACKPROJECTION_W_T;
}

// A plan: host descriptors + the device geometry tables for one angle set.
class Plan {
public:
    Plan(const std::vector<float> &options, const torch::Tensor &angles, int device, bool is2d)
    {
        desc_ = is2d ? unpack2d(options) : unpack3d(options);
        TORCH_CHECK(angles.dim() == 1, "Expected angles shape [V], got ", angles.sizes());
        auto a = angles.detach().to(torch::kCPU, torch::kFloat32).contiguous();
        nview_ = (int)a.numel();
        device_ = device;
        const int rc = xr_plan3d_create(&desc_.vol, &desc_.geo, a.data_ptr<float>(), nview_, device, &plan_);
        if (rc) raise_xr(rc, "xr_plan3d_create");
    }
    ~Plan()
    {
        if (plan_) xr_plan3d_destroy(plan_);
    }
    Plan(const Plan &) = delete;
    Plan &operator=(const Plan &) = delete;

    int num_view() const { return nview_; }
    bool is2d() const { return desc_.is2d; }
    int64_t handle() const { return reinterpret_cast<int64_t>(plan_); }

    std::vector<int64_t> volume_shape(int64_t B, int64_t C) const
    {
        if (desc_.is2d) return {B, C, desc_.H, desc_.W};
        return {B, C, desc_.D, desc_.H, desc_.W};
    }
    std::vector<int64_t> sino_shape(int64_t B, int64_t C) const
    {
        if (desc_.is2d) return {B, C, nview_, desc_.Nu};
        return {B, C, nview_, desc_.Nv, desc_.Nu};
    }

    int64_t workspace_bytes(int op, int64_t nbatch) const
    {
        return (int64_t)xr_plan3d_workspace_bytes(plan_, op, (int)nbatch);
    }

    int kernel_kind(int op) const
    {
        const int k = xr_plan3d_kernel_kind(plan_, op);
        if (k < 0) raise_xr(k, "xr_plan3d_kernel_kind");
        return k;
    }

    std::vector<int> cell_classes() const
    {
        std::vector<int> n(3, 0);
        const int rc = xr_plan3d_cell_classes(plan_, n.data());
        if (rc) raise_xr(rc, "xr_plan3d_cell_classes");
        return n;
    }

    std::vector<int> view_flags() const
    {
        std::vector<int> x(nview_), k(nview_);
        xr_plan3d_view_flags(plan_, x.data(), k.data());
        std::vector<int> out;
        for (int v = 0; v < nview_; ++v) out.push_back(x[v] | (k[v] << 1));
        return out;
    }

    void check_input(int op, const torch::Tensor &x) const
    {
        TORCH_CHECK(x.is_cuda(), "a device plan computes GPU tensors; CPU tensors take a host plan");
        TORCH_CHECK(x.scalar_type() == torch::kFloat32, "input must be float32");
        TORCH_CHECK(x.is_contiguous(), "Expected all tensors to be contiguous");
        TORCH_CHECK(x.get_device() == device_, "input is on device ", x.get_device(), " but the plan on ", device_);
        const int nd = desc_.is2d ? 4 : 5;
        TORCH_CHECK(x.dim() == nd, desc_.is2d ? (volume_in(op) ? "Expected image shape [B, C, H, W]"
                                                               : "Expected projection shape [B, C, V, Nu]")
                                              : (volume_in(op) ? "Expected image shape [B, C, D, H, W]"
                                                               : "Expected projection shape [B, C, V, Nv, Nu]"));
        const auto want = volume_in(op) ? volume_shape(x.size(0), x.size(1)) : sino_shape(x.size(0), x.size(1));
        TORCH_CHECK(x.sizes().vec() == want, volume_in(op) ? "Image shape does not match geometry: "
                                                           : "Projection shape does not match geometry: ",
                    x.sizes(), " vs expected ", want);
    }

    // out and workspace are caller-provided (graph capture / preallocated loops).
    void execute_into(int op, const torch::Tensor &x, torch::Tensor &out, const torch::Tensor &ws, int64_t stream,
                      int tile_rows, int voxel_chunk, int phase) const
    {
        check_input(op, x);
        const int64_t nb = x.size(0) * x.size(1);
        const auto want = volume_in(op) ? sino_shape(x.size(0), x.size(1)) : volume_shape(x.size(0), x.size(1));
        TORCH_CHECK(out.is_cuda() && out.scalar_type() == torch::kFloat32 && out.is_contiguous() &&
                        out.sizes().vec() == want && out.get_device() == device_,
                    "output must be a contiguous float32 GPU tensor of shape ", want);
        const size_t need = xr_plan3d_workspace_bytes(plan_, op, (int)nb);
        void *wsp = nullptr;
        size_t wsb = 0;
        if (ws.defined() && ws.numel() > 0) {
            TORCH_CHECK(ws.is_cuda() && ws.is_contiguous() && ws.get_device() == device_, "workspace must be a contiguous GPU tensor");
            wsp = ws.data_ptr();
            wsb = (size_t)ws.numel() * ws.element_size();
        }
        TORCH_CHECK(wsb >= need, "workspace of ", wsb, " bytes, need ", need);
        const int rc = xr_plan3d_execute_phase(plan_, op, phase, x.data_ptr<float>(), out.data_ptr<float>(), (int)nb,
                                               wsp, wsb, tile_rows, voxel_chunk, reinterpret_cast<void *>(stream));
        if (rc) raise_xr(rc, "xr_plan3d_execute");
    }

    // FDK / FBP in one pass: cosine weight, row filter, weighted backprojection.  `filter`
    // (2 Nu - 1 taps, scaling folded in) and `weight` (Nu) are float32 GPU tensors.
    torch::Tensor fdk(const torch::Tensor &x, const torch::Tensor &filter, const torch::Tensor &weight, int64_t stream,
                      int voxel_chunk) const
    {
        const int op = XR_OP_BACKPROJECTION_W;
        check_input(op, x);
        const int64_t nu = desc_.geo.nu;
        for (const torch::Tensor *t : {&filter, &weight})
            TORCH_CHECK(t->is_cuda() && t->scalar_type() == torch::kFloat32 && t->is_contiguous() &&
                            t->get_device() == device_,
                        "filter and weight must be contiguous float32 GPU tensors on the plan's device");
        TORCH_CHECK(filter.numel() == 2 * nu - 1, "filter must have 2*Nu-1 = ", 2 * nu - 1, " taps, got ",
                    filter.numel());
        TORCH_CHECK(weight.numel() == nu, "weight must have Nu = ", nu, " entries, got ", weight.numel());
        const int64_t B = x.size(0), C = x.size(1);
        auto out = torch::empty(volume_shape(B, C), x.options());
        const int64_t need = workspace_bytes(op, B * C);
        auto ws = torch::empty({need > 0 ? need : 0}, x.options().dtype(torch::kUInt8));
        const int rc = xr_plan3d_execute_fdk(plan_, x.data_ptr<float>(), out.data_ptr<float>(), (int)(B * C),
                                             filter.data_ptr<float>(), weight.data_ptr<float>(),
                                             need > 0 ? ws.data_ptr() : nullptr, (size_t)need, voxel_chunk,
                                             reinterpret_cast<void *>(stream));
        if (rc) raise_xr(rc, "xr_plan3d_execute_fdk");
        return out;
    }

    torch::Tensor execute(int op, const torch::Tensor &x, int64_t stream) const
    {
        check_input(op, x);
        const int64_t B = x.size(0), C = x.size(1);
        const auto shape = volume_in(op) ? sino_shape(B, C) : volume_shape(B, C);
        auto out = torch::empty(shape, x.options());
        const int64_t need = workspace_bytes(op, B * C);
        auto ws = torch::empty({need > 0 ? need : 0}, x.options().dtype(torch::kUInt8));
        execute_into(op, x, out, ws, stream, 0, 0, XR_PHASE_ALL);
        return out;
    }

private:
    Desc desc_;
    int nview_ = 0;
    int device_ = 0;
    xr_plan3d *plan_ = nullptr;
};

// A host plan: the same operators on CPU tensors (xr_host_plan3d_*), the counterpart of
// the reference's CPU dispatch (src/api/torch_binding.cpp:421-425).  Only CPU tensors are
// accepted; a GPU tensor never reaches it.
class HostPlan {
public:
    HostPlan(const std::vector<float> &options, const torch::Tensor &angles, bool is2d)
    {
        desc_ = is2d ? unpack2d(options) : unpack3d(options);
        TORCH_CHECK(angles.dim() == 1, "Expected angles shape [V], got ", angles.sizes());
        auto a = angles.detach().to(torch::kCPU, torch::kFloat32).contiguous();
        nview_ = (int)a.numel();
        const int rc = xr_host_plan3d_create(&desc_.vol, &desc_.geo, a.data_ptr<float>(), nview_, &plan_);
        if (rc) raise_xr(rc, "xr_host_plan3d_create");
    }
    ~HostPlan()
    {
        if (plan_) xr_host_plan3d_destroy(plan_);
    }
    HostPlan(const HostPlan &) = delete;
    HostPlan &operator=(const HostPlan &) = delete;

    int num_view() const { return nview_; }
    bool is2d() const { return desc_.is2d; }

    std::vector<int64_t> volume_shape(int64_t B, int64_t C) const
    {
        if (desc_.is2d) return {B, C, desc_.H, desc_.W};
        return {B, C, desc_.D, desc_.H, desc_.W};
    }
    std::vector<int64_t> sino_shape(int64_t B, int64_t C) const
    {
        if (desc_.is2d) return {B, C, nview_, desc_.Nu};
        return {B, C, nview_, desc_.Nv, desc_.Nu};
    }

    // stream: accepted for the same call signature as Plan::execute, unused on the host.
    torch::Tensor execute(int op, const torch::Tensor &x, int64_t /*stream*/, int64_t num_threads) const
    {
        TORCH_CHECK(!x.is_cuda(), "a host plan computes CPU tensors; GPU tensors take a device plan");
        TORCH_CHECK(x.scalar_type() == torch::kFloat32, "input must be float32");
        TORCH_CHECK(x.is_contiguous(), "Expected all tensors to be contiguous");
        const int nd = desc_.is2d ? 4 : 5;
        TORCH_CHECK(x.dim() == nd, desc_.is2d ? (volume_in(op) ? "Expected image shape [B, C, H, W]"
                                                               : "Expected projection shape [B, C, V, Nu]")
                                              : (volume_in(op) ? "Expected image shape [B, C, D, H, W]"
                                                               : "Expected projection shape [B, C, V, Nv, Nu]"));
        const int64_t B = x.size(0), C = x.size(1);
        const auto want = volume_in(op) ? volume_shape(B, C) : sino_shape(B, C);
        TORCH_CHECK(x.sizes().vec() == want, volume_in(op) ? "Image shape does not match geometry: "
                                                           : "Projection shape does not match geometry: ",
                    x.sizes(), " vs expected ", want);
        auto out = torch::empty(volume_in(op) ? sino_shape(B, C) : volume_shape(B, C), x.options());
        int rc;
        {
            py::gil_scoped_release nogil;
            rc = xr_host_plan3d_execute(plan_, op, x.data_ptr<float>(), out.data_ptr<float>(), (int)(B * C),
                                        (int)num_threads);
        }
        if (rc) raise_xr(rc, "xr_host_plan3d_execute");
        return out;
    }

private:
    Desc desc_;
    int nview_ = 0;
    xr_host_plan3d *plan_ = nullptr;
};

// A batch of plans executed by single launches (xr_batch3d_*).
class Batch {
public:
    Batch(const std::vector<std::shared_ptr<Plan>> &plans, const std::vector<int> &num_batch)
        : plans_(plans), nb_(num_batch)
    {
        TORCH_CHECK(plans.size() == num_batch.size(), "one batch count per plan");
        std::vector<xr_plan3d *> h;
        for (const auto &p : plans) h.push_back(reinterpret_cast<xr_plan3d *>(p->handle()));
        const int rc = xr_batch3d_create(h.data(), (int)h.size(), num_batch.data(), &batch_);
        if (rc) raise_xr(rc, "xr_batch3d_create");
    }
    ~Batch()
    {
        if (batch_) xr_batch3d_destroy(batch_);
    }
    Batch(const Batch &) = delete;
    Batch &operator=(const Batch &) = delete;

    int64_t workspace_bytes(int op) const { return (int64_t)xr_batch3d_workspace_bytes(batch_, op); }

    void execute_into(int op, const std::vector<torch::Tensor> &xs, std::vector<torch::Tensor> &outs,
                      const torch::Tensor &ws, int64_t stream, int tile_rows, int voxel_chunk) const
    {
        TORCH_CHECK(xs.size() == plans_.size() && outs.size() == plans_.size(), "one input and output per plan");
        std::vector<const float *> in;
        std::vector<float *> out;
        for (size_t i = 0; i < xs.size(); ++i) {
            plans_[i]->check_input(op, xs[i]);
            TORCH_CHECK(xs[i].size(0) * xs[i].size(1) == nb_[i], "instance ", i, " batch count differs from the batch");
            const auto want = volume_in(op) ? plans_[i]->sino_shape(xs[i].size(0), xs[i].size(1))
                                            : plans_[i]->volume_shape(xs[i].size(0), xs[i].size(1));
            TORCH_CHECK(outs[i].is_cuda() && outs[i].is_contiguous() && outs[i].scalar_type() == torch::kFloat32 &&
                            outs[i].sizes().vec() == want,
                        "output ", i, " must be a contiguous float32 GPU tensor of shape ", want);
            in.push_back(xs[i].data_ptr<float>());
            out.push_back(outs[i].data_ptr<float>());
        }
        const size_t need = xr_batch3d_workspace_bytes(batch_, op);
        TORCH_CHECK(ws.defined() && ws.is_cuda() && (size_t)ws.numel() * ws.element_size() >= need,
                    "batch workspace needs ", need, " bytes");
        const int rc = xr_batch3d_execute(batch_, op, in.data(), out.data(), ws.data_ptr(),
                                          (size_t)ws.numel() * ws.element_size(), tile_rows, voxel_chunk,
                                          reinterpret_cast<void *>(stream));
        if (rc) raise_xr(rc, "xr_batch3d_execute");
    }

private:
    std::vector<std::shared_ptr<Plan>> plans_;
    std::vector<int> nb_;
    xr_batch3d *batch_ = nullptr;
};

// Plans of the reference-named functions, found without rebuilding tables:
//  * by identity: the (options, angles) tensor objects and their version counters, each
//    validated through a weak reference, so a repeated call with the same tensors does no
//    device->host copy, no synchronisation and no allocation beyond its output and
//    workspace (torch's caching allocator);
//  * by content: the option values, the angle bytes, the device and the dimension, so
//    equal geometry in new tensors reuses the plan after one host copy of both.
// Plans are freed only on eviction (least recently used beyond kMaxPlans), never right
// after a launch that reads their tables.
class PlanCache {
public:
    static constexpr size_t kMaxPlans = 64;

    std::shared_ptr<Plan> get(const torch::Tensor &options, const torch::Tensor &angles, int device, bool is2d)
    {
        std::lock_guard<std::mutex> lk(mu_);
        const std::string ikey = identity_key(options, angles, device, is2d);
        auto it = by_id_.find(ikey);
        if (it != by_id_.end()) {
            IdEntry &e = it->second->second;
            if (e.opts.lock().get() == options.unsafeGetTensorImpl() &&
                e.angs.lock().get() == angles.unsafeGetTensorImpl()) {
                by_id_lru_.splice(by_id_lru_.begin(), by_id_lru_, it->second);
                return e.plan;
            }
            by_id_lru_.erase(it->second);
            by_id_.erase(it);
        }
        const auto o = host_floats(options, is2d ? 12 : 18, "Option");
        auto a = angles.detach().to(torch::kCPU, torch::kFloat32).contiguous();
        std::string vkey(reinterpret_cast<const char *>(o.data()), o.size() * sizeof(float));
        vkey.append(reinterpret_cast<const char *>(a.data_ptr<float>()), (size_t)a.numel() * sizeof(float));
        vkey.append(reinterpret_cast<const char *>(&device), sizeof(device));
        vkey.push_back(is2d ? '2' : '3');
        std::shared_ptr<Plan> plan;
        auto vt = by_val_.find(vkey);
        if (vt != by_val_.end()) {
            by_val_lru_.splice(by_val_lru_.begin(), by_val_lru_, vt->second);
            plan = vt->second->second;
        } else {
            plan = std::make_shared<Plan>(o, a, device, is2d);
            by_val_lru_.emplace_front(vkey, plan);
            by_val_[vkey] = by_val_lru_.begin();
            evict(by_val_, by_val_lru_);
        }
        by_id_lru_.emplace_front(ikey, IdEntry{Weak(options.getIntrusivePtr()), Weak(angles.getIntrusivePtr()), plan});
        by_id_[ikey] = by_id_lru_.begin();
        evict(by_id_, by_id_lru_);
        return plan;
    }

    size_t size() const
    {
        std::lock_guard<std::mutex> lk(mu_);
        return by_val_.size();
    }

    void clear()
    {
        std::lock_guard<std::mutex> lk(mu_);
        by_id_.clear();
        by_id_lru_.clear();
        by_val_.clear();
        by_val_lru_.clear();
    }

private:
    using Weak = c10::weak_intrusive_ptr<c10::TensorImpl, c10::UndefinedTensorImpl>;
    struct IdEntry {
        Weak opts, angs;
        std::shared_ptr<Plan> plan;
    };
    template <class T>
    using Lru = std::list<std::pair<std::string, T>>;
    template <class T>
    using Index = std::unordered_map<std::string, typename Lru<T>::iterator>;

    static std::string identity_key(const torch::Tensor &o, const torch::Tensor &a, int device, bool is2d)
    {
        const int64_t f[6] = {reinterpret_cast<int64_t>(o.unsafeGetTensorImpl()), (int64_t)o._version(),
                              reinterpret_cast<int64_t>(a.unsafeGetTensorImpl()), (int64_t)a._version(),
                              (int64_t)device, (int64_t)is2d};
        return std::string(reinterpret_cast<const char *>(f), sizeof(f));
    }

    template <class T>
    static void evict(Index<T> &index, Lru<T> &lru)
    {
        while (lru.size() > kMaxPlans) {
            index.erase(lru.back().first);
            lru.pop_back();
        }
    }

    mutable std::mutex mu_;
    Lru<IdEntry> by_id_lru_;
    Index<IdEntry> by_id_;
    Lru<std::shared_ptr<Plan>> by_val_lru_;
    Index<std::shared_ptr<Plan>> by_val_;
};

// Host plans of the reference-named functions on CPU tensors, keyed by content (the
// options and angles of a CPU call live in host memory: reading them costs nothing).
class HostPlanCache {
public:
    std::shared_ptr<HostPlan> get(const torch::Tensor &options, const torch::Tensor &angles, bool is2d)
    {
        const auto o = host_floats(options, is2d ? 12 : 18, "Option");
        auto a = angles.detach().to(torch::kCPU, torch::kFloat32).contiguous();
        std::string key(reinterpret_cast<const char *>(o.data()), o.size() * sizeof(float));
        key.append(reinterpret_cast<const char *>(a.data_ptr<float>()), (size_t)a.numel() * sizeof(float));
        key.push_back(is2d ? '2' : '3');
        std::lock_guard<std::mutex> lk(mu_);
        auto it = index_.find(key);
        if (it != index_.end()) {
            lru_.splice(lru_.begin(), lru_, it->second);
            return it->second->second;
        }
        auto plan = std::make_shared<HostPlan>(o, a, is2d);
        lru_.emplace_front(key, plan);
        index_[key] = lru_.begin();
        while (lru_.size() > 16) {
            index_.erase(lru_.back().first);
            lru_.pop_back();
        }
        return plan;
    }

private:
    std::mutex mu_;
    std::list<std::pair<std::string, std::shared_ptr<HostPlan>>> lru_;
    std::unordered_map<std::string, std::list<std::pair<std::string, std::shared_ptr<HostPlan>>>::iterator> index_;
};

HostPlanCache &host_plan_cache()
{
    static HostPlanCache *cache = new HostPlanCache();
    return *cache;
}

PlanCache &plan_cache()
{
    static PlanCache *cache = new PlanCache();  // never destroyed: plans outlive interpreter teardown order
    return *cache;
}

// stream < 0: torch's current stream on `device` (the default of the reference-named
// functions); otherwise the hipStream_t handle as an integer (0 = legacy default stream).
int64_t resolve_stream(int64_t stream, int device)
{
    if (stream >= 0) return stream;
    return reinterpret_cast<int64_t>(c10::hip::getCurrentHIPStream((c10::DeviceIndex)device).stream());
}

// Reference-named operator functions: (input, options, angles[, stream]) -> output, as in
// src/api/torch_binding.cpp:407-523 (3D) and :139-351 (2D).  The plan comes from the
// cache above; the output and workspace from torch's caching allocator; the work goes to
// torch's current stream unless a stream is given.
torch::Tensor run_op(int op, bool is2d, const torch::Tensor &x, const torch::Tensor &options,
                     const torch::Tensor &angles, int64_t stream)
{
    if (!x.is_cuda()) {
        // CPU tensors: the host executor, as the reference runs its CPU kernels
        TORCH_CHECK(!options.is_cuda() && !angles.is_cuda(), "Expected all tensors on the same device");
        TORCH_CHECK(options.scalar_type() == torch::kFloat32, "options must be float32");
        TORCH_CHECK(angles.scalar_type() == torch::kFloat32, "ang must be float32");
        TORCH_CHECK(angles.is_contiguous() && options.is_contiguous(), "Expected all tensors to be contiguous");
        TORCH_CHECK(angles.dim() == 1, "Expected angles shape [V], got ", angles.sizes());
        return host_plan_cache().get(options, angles, is2d)->execute(op, x, 0, 0);
    }
    TORCH_CHECK(!angles.is_cuda() || angles.get_device() == x.get_device(),
                "Expected all tensors on the same device");
    TORCH_CHECK(!options.is_cuda() || options.get_device() == x.get_device(),
                "Expected all tensors on the same device");
    TORCH_CHECK(options.scalar_type() == torch::kFloat32, "options must be float32");
    TORCH_CHECK(angles.scalar_type() == torch::kFloat32, "ang must be float32");
    TORCH_CHECK(angles.is_contiguous() && options.is_contiguous(), "Expected all tensors to be contiguous");
    TORCH_CHECK(angles.dim() == 1, "Expected angles shape [V], got ", angles.sizes());
    const int dev = x.get_device();
    auto plan = plan_cache().get(options, angles, dev, is2d);
    return plan->execute(op, x, resolve_stream(stream, dev));
}

// system_matrix2d / system_matrix3d (src/api/torch_binding.cpp:353-405): CPU tensors in,
// coalesced sparse COO [V*Nv*Nu, D*H*W] out.
torch::Tensor system_matrix(bool is2d, const torch::Tensor &options, const torch::Tensor &ang)
{
    TORCH_CHECK(!options.is_cuda() && !ang.is_cuda(), is2d ? "system_matrix2d builds the matrix on CPU"
                                                           : "system_matrix3d builds the matrix on CPU");
    TORCH_CHECK(options.scalar_type() == torch::kFloat32, "options must be float32");
    TORCH_CHECK(ang.scalar_type() == torch::kFloat32, "ang must be float32");
    TORCH_CHECK(options.is_contiguous() && ang.is_contiguous(), "Expected all tensors to be contiguous");
    const auto o = host_floats(options, is2d ? 12 : 18, "Option");
    const int V = (int)ang.numel();
    xr_coo *coo = nullptr;
    int rc;
    int64_t nrows, ncols;
    if (is2d) {
        const int type = (int)o[11];
        xr_volume2d v2{(int)o[0], (int)o[1], o[3], o[4], o[8], o[9]};
        xr_geometry2d g2{type, (int)o[2], o[5], o[10], o[6], o[7]};
        rc = xr_system_matrix2d(&v2, &g2, ang.data_ptr<float>(), V, &coo);
        nrows = (int64_t)V * g2.nu;
        ncols = (int64_t)v2.nx * v2.ny;
    } else {
        const Desc d = unpack3d(o);
        rc = xr_system_matrix3d(&d.vol, &d.geo, ang.data_ptr<float>(), V, &coo);
        nrows = (int64_t)V * d.Nv * d.Nu;
        ncols = (int64_t)d.D * d.H * d.W;
    }
    if (rc) raise_xr(rc, is2d ? "xr_system_matrix2d" : "xr_system_matrix3d");
    const int64_t n = xr_coo_nnz(coo);
    auto rows = torch::empty({n}, torch::kInt64), cols = torch::empty({n}, torch::kInt64);
    auto vals = torch::empty({n}, torch::kFloat32);
    xr_coo_copy(coo, rows.data_ptr<int64_t>(), cols.data_ptr<int64_t>(), vals.data_ptr<float>());
    xr_coo_destroy(coo);
    return torch::sparse_coo_tensor(torch::stack({rows, cols}), vals, {nrows, ncols}).coalesce();
}

// The same triplets emitted on a GPU (xr_system_matrix*_gpu): rows, cols (int64) and
// values (float32) as device tensors in the host builder's order, before coalescing.
std::vector<torch::Tensor> system_matrix_triplets_device(bool is2d, const torch::Tensor &options,
                                                         const torch::Tensor &ang, int64_t device)
{
    TORCH_CHECK(options.scalar_type() == torch::kFloat32, "options must be float32");
    TORCH_CHECK(ang.scalar_type() == torch::kFloat32, "ang must be float32");
    const auto o = host_floats(options, is2d ? 12 : 18, "Option");
    auto a = ang.detach().to(torch::kCPU, torch::kFloat32).contiguous();
    const int V = (int)a.numel();
    const int dev = (int)device;
    const int64_t stream = resolve_stream(-1, dev);
    xr_sysmat_gpu *m = nullptr;
    int rc;
    if (is2d) {
        xr_volume2d v2{(int)o[0], (int)o[1], o[3], o[4], o[8], o[9]};
        xr_geometry2d g2{(int)o[11], (int)o[2], o[5], o[10], o[6], o[7]};
        rc = xr_system_matrix2d_gpu(&v2, &g2, a.data_ptr<float>(), V, dev, reinterpret_cast<void *>(stream), &m);
    } else {
        const Desc d = unpack3d(o);
        rc = xr_system_matrix3d_gpu(&d.vol, &d.geo, a.data_ptr<float>(), V, dev, reinterpret_cast<void *>(stream), &m);
    }
    if (rc) raise_xr(rc, is2d ? "xr_system_matrix2d_gpu" : "xr_system_matrix3d_gpu");
    std::unique_ptr<xr_sysmat_gpu, void (*)(xr_sysmat_gpu *)> guard(m, xr_sysmat_gpu_destroy);
    const int64_t n = xr_sysmat_gpu_nnz(m);
    const auto dopt = torch::TensorOptions().device(torch::kCUDA, dev);
    auto rows = torch::empty({n}, dopt.dtype(torch::kInt64)), cols = torch::empty({n}, dopt.dtype(torch::kInt64));
    auto vals = torch::empty({n}, dopt.dtype(torch::kFloat32));
    rc = xr_sysmat_gpu_emit(m, rows.data_ptr<int64_t>(), cols.data_ptr<int64_t>(), vals.data_ptr<float>(),
                            reinterpret_cast<void *>(stream));
    if (rc) raise_xr(rc, "xr_sysmat_gpu_emit");
    // the tables and counts the emission reads are freed by `guard`: wait for it first
    TORCH_CHECK(hipStreamSynchronize(reinterpret_cast<hipStream_t>(stream)) == hipSuccess,
                "stream synchronisation after the system-matrix emission failed");
    return {rows, cols, vals};
}

// Host builder triplets, uncoalesced, in traversal order (tests compare them with the
// GPU emission).
std::vector<torch::Tensor> system_matrix_triplets_host(bool is2d, const torch::Tensor &options,
                                                       const torch::Tensor &ang)
{
    const auto o = host_floats(options, is2d ? 12 : 18, "Option");
    auto a = ang.detach().to(torch::kCPU, torch::kFloat32).contiguous();
    const int V = (int)a.numel();
    xr_coo *coo = nullptr;
    int rc;
    if (is2d) {
        xr_volume2d v2{(int)o[0], (int)o[1], o[3], o[4], o[8], o[9]};
        xr_geometry2d g2{(int)o[11], (int)o[2], o[5], o[10], o[6], o[7]};
        rc = xr_system_matrix2d(&v2, &g2, a.data_ptr<float>(), V, &coo);
    } else {
        const Desc d = unpack3d(o);
        rc = xr_system_matrix3d(&d.vol, &d.geo, a.data_ptr<float>(), V, &coo);
    }
    if (rc) raise_xr(rc, is2d ? "xr_system_matrix2d" : "xr_system_matrix3d");
    const int64_t n = xr_coo_nnz(coo);
    auto rows = torch::empty({n}, torch::kInt64), cols = torch::empty({n}, torch::kInt64);
    auto vals = torch::empty({n}, torch::kFloat32);
    xr_coo_copy(coo, rows.data_ptr<int64_t>(), cols.data_ptr<int64_t>(), vals.data_ptr<float>());
    xr_coo_destroy(coo);
    return {rows, cols, vals};
}

#define XR_DEF_OP(name, op, is2d)                                                                         \
    m.def(                                                                                                   \
        name,                                                                                                \
        [](const torch::Tensor &x, const torch::Tensor &o, const torch::Tensor &a, int64_t stream) {         \
            return run_op(op, is2d, x, o, a, stream);                                                        \
        },                                                                                                   \
        py::arg("input"), py::arg("options"), py::arg("angles"), py::arg("stream") = -1)

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m)
{
    m.doc() = "xrecon_amd native binding over libxrecon_amd.so (gfx950)";
    // dynamic_attr: per-plan Python state (the autotuner keeps its choices on the plan).
    py::class_<Plan, std::shared_ptr<Plan>>(m, "Plan", py::dynamic_attr())
        .def(py::init([](const std::vector<float> &options, const torch::Tensor &angles, int device, bool is2d) {
                 return std::make_shared<Plan>(options, angles, device, is2d);
             }),
             py::arg("options"), py::arg("angles"), py::arg("device"), py::arg("is2d") = false)
        .def_property_readonly("num_view", &Plan::num_view)
        .def_property_readonly("is2d", &Plan::is2d)
        .def_property_readonly("handle", &Plan::handle)
        .def("workspace_bytes", &Plan::workspace_bytes)
        .def("view_flags", &Plan::view_flags)
        .def("cell_classes", &Plan::cell_classes)
        .def("kernel_kind", &Plan::kernel_kind)
        .def("execute", &Plan::execute, py::arg("op"), py::arg("input"), py::arg("stream") = 0)
        .def("fdk", &Plan::fdk, py::arg("input"), py::arg("filter"), py::arg("weight"), py::arg("stream") = 0,
             py::arg("voxel_chunk") = 0)
        .def("execute_into", &Plan::execute_into, py::arg("op"), py::arg("input"), py::arg("out"),
             py::arg("workspace"), py::arg("stream") = 0, py::arg("tile_rows") = 0, py::arg("voxel_chunk") = 0,
             py::arg("phase") = XR_PHASE_ALL);

    py::class_<HostPlan, std::shared_ptr<HostPlan>>(m, "HostPlan", py::dynamic_attr())
        .def(py::init([](const std::vector<float> &options, const torch::Tensor &angles, bool is2d) {
                 return std::make_shared<HostPlan>(options, angles, is2d);
             }),
             py::arg("options"), py::arg("angles"), py::arg("is2d") = false)
        .def_property_readonly("num_view", &HostPlan::num_view)
        .def_property_readonly("is2d", &HostPlan::is2d)
        .def("execute", &HostPlan::execute, py::arg("op"), py::arg("input"), py::arg("stream") = 0,
             py::arg("num_threads") = 0);

    py::class_<Batch, std::shared_ptr<Batch>>(m, "Batch")
        .def(py::init([](const std::vector<std::shared_ptr<Plan>> &plans, const std::vector<int> &num_batch) {
                 return std::make_shared<Batch>(plans, num_batch);
             }),
             py::arg("plans"), py::arg("num_batch"))
        .def("workspace_bytes", &Batch::workspace_bytes)
        .def("execute_into", &Batch::execute_into, py::arg("op"), py::arg("inputs"), py::arg("outputs"),
             py::arg("workspace"), py::arg("stream") = 0, py::arg("tile_rows") = 0, py::arg("voxel_chunk") = 0);
    m.attr("BATCH_MAX") = XR_BATCH_MAX;
    m.attr("FDK_MAX_NU") = 1024;

    m.attr("OP_PROJECTION") = XR_OP_PROJECTION;
    m.attr("OP_PROJECTION_T") = XR_OP_PROJECTION_T;
    m.attr("OP_BACKPROJECTION") = XR_OP_BACKPROJECTION;
    m.attr("OP_BACKPROJECTION_W") = XR_OP_BACKPROJECTION_W;
    m.attr("OP_BACKPROJECTION_T") = XR_OP_BACKPROJECTION_T;
    m.attr("OP_BACKPROJECTION_W_T") = XR_OP_BACKPROJECTION_W_T;
    m.def("version", []() { return std::string(xr_version()); });
    m.def("plan_cache_size", []() { return (int64_t)plan_cache().size(); },
          "number of distinct geometries cached by the reference-named functions");
    m.def("plan_cache_clear", []() { plan_cache().clear(); });

    m.def("system_matrix2d", [](const torch::Tensor &o, const torch::Tensor &a) { return system_matrix(true, o, a); },
          "2D distance-driven projection matrix");
    m.def("system_matrix3d", [](const torch::Tensor &o, const torch::Tensor &a) { return system_matrix(false, o, a); },
          "3D distance-driven projection matrix");
    m.def("system_matrix_triplets",
          [](const torch::Tensor &o, const torch::Tensor &a, bool is2d, int64_t device) {
              return device < 0 ? system_matrix_triplets_host(is2d, o, a)
                                : system_matrix_triplets_device(is2d, o, a, device);
          },
          py::arg("options"), py::arg("angles"), py::arg("is2d"), py::arg("device") = -1,
          "system-matrix triplets (rows, cols, values) before coalescing: host builder (device < 0) or GPU emission");
    XR_DEF_OP("forward_projection2d", XR_OP_PROJECTION, true);
    XR_DEF_OP("backward_projection2d", XR_OP_BACKPROJECTION, true);
    XR_DEF_OP("weighted_backward_projection2d", XR_OP_BACKPROJECTION_W, true);
    XR_DEF_OP("projection_transpose2d", XR_OP_PROJECTION_T, true);
    XR_DEF_OP("backward_projection_transpose2d", XR_OP_BACKPROJECTION_T, true);
    XR_DEF_OP("weighted_backward_projection_transpose2d", XR_OP_BACKPROJECTION_W_T, true);
    XR_DEF_OP("back_projection2d", XR_OP_BACKPROJECTION, true);
    XR_DEF_OP("weighted_back_projection2d", XR_OP_BACKPROJECTION_W, true);
    XR_DEF_OP("forward_projection3d", XR_OP_PROJECTION, false);
    XR_DEF_OP("projection_transpose3d", XR_OP_PROJECTION_T, false);
    XR_DEF_OP("backward_projection3d", XR_OP_BACKPROJECTION, false);
    XR_DEF_OP("weighted_backward_projection3d", XR_OP_BACKPROJECTION_W, false);
    XR_DEF_OP("backward_projection_transpose3d", XR_OP_BACKPROJECTION_T, false);
    XR_DEF_OP("weighted_backward_projection_transpose3d", XR_OP_BACKPROJECTION_W_T, false);
}
