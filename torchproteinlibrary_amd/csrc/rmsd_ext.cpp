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
}
