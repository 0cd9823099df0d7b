// rmsd_ext.cpp -- pybind11 module `_RMSD`: the tensor-facing glue over the C
// ABI in include/tpl_rmsd.h.
//
// Mirrors the reference module of the same name (Layers/RMSD/main.cpp:4-7) and
// its checks (Layers/RMSD/Coords2RMSD/coords2rmsd_interface.cpp:58-65,
// Math/nUtil.h:49-57): device, contiguity and dtype checks raise RuntimeError,
// and the ndim check that the reference's no-op ERROR() skips is enforced.
// Work goes to torch's current HIP stream of the input's device.
#include <torch/extension.h>
#include <c10/hip/HIPStream.h>

#include <tuple>
#include <vector>

#include "tpl_rmsd.h"

namespace {

#define RMSD_CHECK_GPU(x)                                                     \
    TORCH_CHECK((x).is_cuda(), #x " must be a GPU tensor");                   \
    TORCH_CHECK((x).is_contiguous(), #x " must be contiguous")

void check_rc(int rc, const char* what) {
    TORCH_CHECK(rc == TPL_RMSD_OK, what, " failed: ", tpl_rmsd_strerror(rc));
}

int dtype_code(const torch::Tensor& t) {
    if (t.scalar_type() == torch::kFloat) return TPL_RMSD_F32;
    if (t.scalar_type() == torch::kDouble) return TPL_RMSD_F64;
    TORCH_CHECK(false, "Coords2RMSD: coordinates must be float32 or float64, got ", t.scalar_type());
    return -1;
}

// input, target [B, 3*A] of one dtype and device, num_atoms int32 [B]; returns (B, A).
std::pair<int, int> check_inputs(const torch::Tensor& input, const torch::Tensor& target,
                                 const torch::Tensor& num_atoms) {
    RMSD_CHECK_GPU(input);
    RMSD_CHECK_GPU(target);
    RMSD_CHECK_GPU(num_atoms);
    TORCH_CHECK(num_atoms.scalar_type() == torch::kInt, "num_atoms must be int32");
    TORCH_CHECK(input.dim() == 2 && input.size(1) % 3 == 0,
                "Coords2RMSD: input must be [batch, 3*num_atoms], got ", input.sizes());
    TORCH_CHECK(target.sizes() == input.sizes(), "Coords2RMSD: target has shape ", target.sizes(),
                ", input ", input.sizes());
    TORCH_CHECK(target.scalar_type() == input.scalar_type(), "target must have the dtype of input");
    TORCH_CHECK(num_atoms.dim() == 1 && num_atoms.size(0) == input.size(0),
                "num_atoms must be [batch]");
    TORCH_CHECK(target.device() == input.device() && num_atoms.device() == input.device(),
                "input, target and num_atoms must be on one device");
    TORCH_CHECK(input.size(0) <= INT32_MAX && input.size(1) / 3 <= INT32_MAX, "input too large");
    return {(int)input.size(0), (int)(input.size(1) / 3)};
}

void check_like(const torch::Tensor& t, const torch::Tensor& like, std::vector<int64_t> shape,
                const char* name) {
    RMSD_CHECK_GPU(t);
    TORCH_CHECK(t.scalar_type() == like.scalar_type(), name, " must have the dtype of input");
    TORCH_CHECK(t.device() == like.device(), name, " must be on the device of input");
    TORCH_CHECK(t.sizes() == torch::IntArrayRef(shape), name, " has shape ", t.sizes());
}

void* state_ptr(const torch::Tensor& t, int B, const torch::Tensor& like) {
    RMSD_CHECK_GPU(t);
    TORCH_CHECK(t.device() == like.device(), "state must be on the device of input");
    const size_t need = tpl_rmsd_state_bytes(B);
    TORCH_CHECK((size_t)t.nbytes() >= need, "state needs ", need, " bytes, has ", t.nbytes());
    return need ? t.data_ptr() : nullptr;
}

uint32_t* status_ptr(const c10::optional<torch::Tensor>& status, const torch::Tensor& like) {
    if (!status.has_value() || !status->defined()) return nullptr;
    RMSD_CHECK_GPU(*status);
    TORCH_CHECK(status->scalar_type() == torch::kInt && status->numel() >= 1,
                "status must be an int32 tensor");
    TORCH_CHECK(status->device() == like.device(), "status must be on the device of input");
    return reinterpret_cast<uint32_t*>(status->data_ptr());
}

hipStream_t stream_of(const torch::Tensor& t) {
    return c10::hip::getCurrentHIPStream(t.device().index()).stream();
}

// Coords2RMSDGPU_forward (reference coords2rmsd_interface.cpp:52-57):
// (centered_coords_src, centered_coords_dst, output, num_atoms, UT) plus the
// forward->backward state (None: no backward follows) and an optional status.
// The inputs need not be centred.
void forward(torch::Tensor input, torch::Tensor target, torch::Tensor output,
             torch::Tensor num_atoms, torch::Tensor UT, c10::optional<torch::Tensor> state,
             c10::optional<torch::Tensor> status) {
    auto [B, A] = check_inputs(input, target, num_atoms);
    check_like(output, input, {B}, "output");
    check_like(UT, input, {B, 3, 3}, "UT");
    c10::DeviceGuard guard(input.device());
    void* st = (state.has_value() && state->defined()) ? state_ptr(*state, B, input) : nullptr;
    check_rc(tpl_rmsd_forward(dtype_code(input), input.data_ptr(), target.data_ptr(),
                              num_atoms.data_ptr<int32_t>(), output.data_ptr(), UT.data_ptr(), st,
                              status_ptr(status, input), B, A, stream_of(input)),
             "Coords2RMSDGPU_forward");
}

// The backward the reference computes in Python (Coords2RMSD.py:40-62):
// gradInput from gradRMSD, both inputs and the forward's state.
void backward(torch::Tensor grad_input, torch::Tensor grad_rmsd, torch::Tensor input,
              torch::Tensor target, torch::Tensor num_atoms, torch::Tensor state,
              c10::optional<torch::Tensor> status) {
    auto [B, A] = check_inputs(input, target, num_atoms);
    check_like(grad_input, input, {B, 3 * (int64_t)A}, "gradInput");
    check_like(grad_rmsd, input, {B}, "gradRMSD");
    c10::DeviceGuard guard(input.device());
    check_rc(tpl_rmsd_backward(dtype_code(input), grad_input.data_ptr(), grad_rmsd.data_ptr(),
                               input.data_ptr(), target.data_ptr(), num_atoms.data_ptr<int32_t>(),
                               state_ptr(state, B, input), status_ptr(status, input), B, A,
                               stream_of(input)),
             "Coords2RMSDGPU_backward");
}

// Timing loop for bench.py: sets of [input, target, num_atoms, rmsd, UT,
// state, grad_rmsd, grad_input], rotated per step.  Returns HIP-event ms for
// `steps` back-to-back forward+backward steps, and the average ms of one
// forward and of one backward launched `reps` times alone.
std::tuple<double, double, double> bench(std::vector<std::vector<torch::Tensor>> sets,
                                         int64_t steps, int64_t warmup, int64_t reps) {
    TORCH_CHECK(!sets.empty() && steps > 0 && warmup >= 0 && reps > 0, "bench: bad arguments");
    struct Set {
        int dt, B, A;
        const void *x, *y, *gr;
        const int32_t* n;
        void *r, *ut, *st, *gx;
    };
    std::vector<Set> ss;
    for (auto& t : sets) {
        TORCH_CHECK(t.size() == 8, "bench: a set has 8 tensors");
        auto [B, A] = check_inputs(t[0], t[1], t[2]);
        check_like(t[3], t[0], {B}, "output");
        check_like(t[4], t[0], {B, 3, 3}, "UT");
        check_like(t[6], t[0], {B}, "gradRMSD");
        check_like(t[7], t[0], {B, 3 * (int64_t)A}, "gradInput");
        ss.push_back({dtype_code(t[0]), B, A, t[0].data_ptr(), t[1].data_ptr(), t[6].data_ptr(),
                      t[2].data_ptr<int32_t>(), t[3].data_ptr(), t[4].data_ptr(),
                      state_ptr(t[5], B, t[0]), t[7].data_ptr()});
    }
    c10::DeviceGuard guard(sets[0][0].device());
    hipStream_t stream = stream_of(sets[0][0]);
    auto fwd = [&](const Set& q) {
        check_rc(tpl_rmsd_forward(q.dt, q.x, q.y, q.n, q.r, q.ut, q.st, nullptr, q.B, q.A, stream),
                 "tpl_rmsd_forward");
    };
    auto bwd = [&](const Set& q) {
        check_rc(tpl_rmsd_backward(q.dt, q.gx, q.gr, q.x, q.y, q.n, q.st, nullptr, q.B, q.A, stream),
                 "tpl_rmsd_backward");
    };
    hipEvent_t e0, e1;
    TORCH_CHECK(hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess,
                "hipEventCreate failed");
    auto timed = [&](auto&& body, int64_t n) {
        TORCH_CHECK(hipEventRecord(e0, stream) == hipSuccess, "hipEventRecord failed");
        for (int64_t i = 0; i < n; ++i) body(i);
        TORCH_CHECK(hipEventRecord(e1, stream) == hipSuccess, "hipEventRecord failed");
        TORCH_CHECK(hipEventSynchronize(e1) == hipSuccess, "hipEventSynchronize failed");
        float ms = 0.f;
        TORCH_CHECK(hipEventElapsedTime(&ms, e0, e1) == hipSuccess, "hipEventElapsedTime failed");
        return (double)ms;
    };
    const size_t S = ss.size();
    auto step = [&](int64_t i) {
        fwd(ss[i % S]);
        bwd(ss[i % S]);
    };
    for (int64_t i = 0; i < warmup; ++i) step(i);
    const double step_ms = timed(step, steps);
    const double f_ms = timed([&](int64_t i) { fwd(ss[i % S]); }, reps) / reps;
    const double b_ms = timed([&](int64_t i) { bwd(ss[i % S]); }, reps) / reps;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return {step_ms, f_ms, b_ms};
}

int64_t state_bytes(int64_t batch) {
    TORCH_CHECK(batch >= 0 && batch <= INT32_MAX, "bad batch");
    return (int64_t)tpl_rmsd_state_bytes((int)batch);
}

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "Coords2RMSD for MI355X: tensor glue over the tpl_rmsd C ABI";
    m.def("Coords2RMSDGPU_forward", &forward, "Coords2RMSD forward on GPU (centring, moments, eigenproblem)");
    m.def("Coords2RMSDGPU_backward", &backward, "Coords2RMSD backward on GPU");
    m.def("state_bytes", &state_bytes, "forward->backward state bytes for a batch");
    m.def("bench", &bench, "time fwd+bwd steps and each kernel alone (event ms)");
}
