// xform_ext.cpp -- pybind11 module `_FullAtomModel` (the coordinate transform
// part): the tensor-facing glue over the C ABI in include/tpl_xform.h.
//
// Mirrors the reference's GPU entry points of the same names and argument
// orders (Layers/FullAtomModel/main.cpp:23-36,
// Layers/FullAtomModel/CoordsTransform/coordsTransformGPU_interface.cpp:5-146)
// and their checks (Math/nUtil.h:55-57): device, contiguity and dtype checks
// raise RuntimeError; the ndim checks the reference's no-op ERROR() skips are
// enforced.  Work goes to torch's current HIP stream of the input's device.
#include <torch/extension.h>
#include <c10/hip/HIPStream.h>

#include <tuple>
#include <vector>

#include "tpl_xform.h"

namespace {

#define XF_CHECK_GPU(x)                                                       \
    TORCH_CHECK((x).is_cuda(), #x " must be a GPU tensor");                   \
    TORCH_CHECK((x).is_contiguous(), #x " must be contiguous")

void check_rc(int rc, const char* what) {
    TORCH_CHECK(rc == TPL_XFORM_OK, what, " failed: ", tpl_xform_strerror(rc));
}

int dtype_code(const torch::Tensor& t) {
    if (t.scalar_type() == torch::kFloat) return TPL_XFORM_F32;
    if (t.scalar_type() == torch::kDouble) return TPL_XFORM_F64;
    TORCH_CHECK(false, "coordinates must be float32 or float64, got ", t.scalar_type());
    return -1;
}

// coords [B, 3A] and num_atoms int32 [B]; returns (B, A).
std::pair<int, int> check_coords(const torch::Tensor& c, const torch::Tensor& num_atoms) {
    XF_CHECK_GPU(c);
    XF_CHECK_GPU(num_atoms);
    TORCH_CHECK(num_atoms.scalar_type() == torch::kInt, "num_atoms must be int32");
    TORCH_CHECK(c.dim() == 2 && c.size(1) % 3 == 0, "coordinates must be [batch, 3*num_atoms], got ",
                c.sizes());
    TORCH_CHECK(num_atoms.dim() == 1 && num_atoms.size(0) == c.size(0), "num_atoms must be [batch]");
    TORCH_CHECK(num_atoms.device() == c.device(), "num_atoms must be on the device of the coordinates");
    TORCH_CHECK(c.size(0) <= INT32_MAX && c.size(1) / 3 <= INT32_MAX, "coordinates too large");
    return {(int)c.size(0), (int)(c.size(1) / 3)};
}

void check_like(const torch::Tensor& t, const torch::Tensor& like, std::vector<int64_t> shape,
                const char* name) {
    XF_CHECK_GPU(t);
    TORCH_CHECK(t.scalar_type() == like.scalar_type(), name, " must have the dtype of the coordinates");
    TORCH_CHECK(t.device() == like.device(), name, " must be on the device of the coordinates");
    TORCH_CHECK(t.sizes() == torch::IntArrayRef(shape), name, " has shape ", t.sizes());
}

uint32_t* status_ptr(const c10::optional<torch::Tensor>& status, const torch::Tensor& like) {
    if (!status.has_value() || !status->defined()) return nullptr;
    XF_CHECK_GPU(*status);
    TORCH_CHECK(status->scalar_type() == torch::kInt && status->numel() >= 1,
                "status must be an int32 tensor");
    TORCH_CHECK(status->device() == like.device(), "status must be on the device of the coordinates");
    return reinterpret_cast<uint32_t*>(status->data_ptr());
}

hipStream_t stream_of(const torch::Tensor& t) {
    return c10::hip::getCurrentHIPStream(t.device().index()).stream();
}

void translate_forward(torch::Tensor input_coords, torch::Tensor output_coords, torch::Tensor T,
                       torch::Tensor num_atoms, c10::optional<torch::Tensor> status) {
    auto [B, A] = check_coords(input_coords, num_atoms);
    check_like(output_coords, input_coords, {B, 3 * (int64_t)A}, "output_coords");
    check_like(T, input_coords, {B, 3}, "T");
    c10::DeviceGuard guard(input_coords.device());
    check_rc(tpl_xform_translate(dtype_code(input_coords), output_coords.data_ptr(),
                                 input_coords.data_ptr(), T.data_ptr(), num_atoms.data_ptr<int32_t>(),
                                 status_ptr(status, input_coords), B, A, stream_of(input_coords)),
             "CoordsTranslateGPU_forward");
}

void translate_backward(torch::Tensor grad_output_coords, torch::Tensor grad_input_coords,
                        torch::Tensor T, torch::Tensor num_atoms,
                        c10::optional<torch::Tensor> status) {
    auto [B, A] = check_coords(grad_output_coords, num_atoms);
    check_like(grad_input_coords, grad_output_coords, {B, 3 * (int64_t)A}, "grad_input_coords");
    (void)T;  // the translation gets no gradient (reference CoordsTransform.py:91)
    c10::DeviceGuard guard(grad_output_coords.device());
    check_rc(tpl_xform_translate_backward(dtype_code(grad_output_coords), grad_input_coords.data_ptr(),
                                          grad_output_coords.data_ptr(), num_atoms.data_ptr<int32_t>(),
                                          status_ptr(status, grad_output_coords), B, A,
                                          stream_of(grad_output_coords)),
             "CoordsTranslateGPU_backward");
}

void rotate(torch::Tensor in, torch::Tensor out, torch::Tensor R, torch::Tensor num_atoms,
            c10::optional<torch::Tensor> status, int transpose, const char* what) {
    auto [B, A] = check_coords(in, num_atoms);
    check_like(out, in, {B, 3 * (int64_t)A}, "output");
    check_like(R, in, {B, 3, 3}, "R");
    c10::DeviceGuard guard(in.device());
    check_rc(tpl_xform_rotate(dtype_code(in), out.data_ptr(), in.data_ptr(), R.data_ptr(), transpose,
                              num_atoms.data_ptr<int32_t>(), status_ptr(status, in), B, A,
                              stream_of(in)),
             what);
}

void rotate_forward(torch::Tensor input_coords, torch::Tensor output_coords, torch::Tensor R,
                    torch::Tensor num_atoms, c10::optional<torch::Tensor> status) {
    rotate(input_coords, output_coords, R, num_atoms, status, 0, "CoordsRotateGPU_forward");
}

void rotate_backward(torch::Tensor grad_output_coords, torch::Tensor grad_input_coords,
                     torch::Tensor R, torch::Tensor num_atoms, c10::optional<torch::Tensor> status) {
    rotate(grad_output_coords, grad_input_coords, R, num_atoms, status, 1, "CoordsRotateGPU_backward");
}

void center_forward(torch::Tensor input_coords, torch::Tensor output_center,
                    torch::Tensor num_atoms, c10::optional<torch::Tensor> status) {
    auto [B, A] = check_coords(input_coords, num_atoms);
    check_like(output_center, input_coords, {B, 3}, "output_center");
    c10::DeviceGuard guard(input_coords.device());
    check_rc(tpl_xform_center(dtype_code(input_coords), output_center.data_ptr(),
                              input_coords.data_ptr(), num_atoms.data_ptr<int32_t>(),
                              status_ptr(status, input_coords), B, A, stream_of(input_coords)),
             "Coords2CenterGPU_forward");
}

void center_backward(torch::Tensor grad_output_center, torch::Tensor grad_input_coords,
                     torch::Tensor num_atoms, c10::optional<torch::Tensor> status) {
    auto [B, A] = check_coords(grad_input_coords, num_atoms);
    check_like(grad_output_center, grad_input_coords, {B, 3}, "grad_output_center");
    c10::DeviceGuard guard(grad_input_coords.device());
    check_rc(tpl_xform_center_backward(dtype_code(grad_input_coords), grad_input_coords.data_ptr(),
                                       grad_output_center.data_ptr(), num_atoms.data_ptr<int32_t>(),
                                       status_ptr(status, grad_input_coords), B, A,
                                       stream_of(grad_input_coords)),
             "Coords2CenterGPU_backward");
}

// Timing loop for bench.py: sets of [coords, out, T, R, center, num_atoms],
// rotated per launch.  Returns the average HIP-event ms of one launch of
// translate, rotate, centre and centre-backward, each run `reps` times.
std::tuple<double, double, double, double> bench(std::vector<std::vector<torch::Tensor>> sets,
                                                 int64_t reps) {
    TORCH_CHECK(!sets.empty() && reps > 0, "bench: bad arguments");
    for (auto& t : sets) {
        TORCH_CHECK(t.size() == 6, "bench: a set has 6 tensors");
        auto [B, A] = check_coords(t[0], t[5]);
        check_like(t[1], t[0], {B, 3 * (int64_t)A}, "out");
        check_like(t[2], t[0], {B, 3}, "T");
        check_like(t[3], t[0], {B, 3, 3}, "R");
        check_like(t[4], t[0], {B, 3}, "center");
    }
    c10::DeviceGuard guard(sets[0][0].device());
    hipStream_t stream = stream_of(sets[0][0]);
    const int dt = dtype_code(sets[0][0]);
    const int B = (int)sets[0][0].size(0), A = (int)(sets[0][0].size(1) / 3);
    hipEvent_t e0, e1;
    TORCH_CHECK(hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess,
                "hipEventCreate failed");
    auto timed = [&](auto&& body) {
        for (int64_t i = 0; i < 5; ++i) body(sets[i % sets.size()]);
        TORCH_CHECK(hipEventRecord(e0, stream) == hipSuccess, "hipEventRecord failed");
        for (int64_t i = 0; i < reps; ++i) body(sets[i % sets.size()]);
        TORCH_CHECK(hipEventRecord(e1, stream) == hipSuccess, "hipEventRecord failed");
        TORCH_CHECK(hipEventSynchronize(e1) == hipSuccess, "hipEventSynchronize failed");
        float ms = 0.f;
        TORCH_CHECK(hipEventElapsedTime(&ms, e0, e1) == hipSuccess, "hipEventElapsedTime failed");
        return (double)ms / reps;
    };
    auto n = [](std::vector<torch::Tensor>& t) { return t[5].data_ptr<int32_t>(); };
    const double tr = timed([&](std::vector<torch::Tensor>& t) {
        check_rc(tpl_xform_translate(dt, t[1].data_ptr(), t[0].data_ptr(), t[2].data_ptr(), n(t),
                                     nullptr, B, A, stream), "translate");
    });
    const double ro = timed([&](std::vector<torch::Tensor>& t) {
        check_rc(tpl_xform_rotate(dt, t[1].data_ptr(), t[0].data_ptr(), t[3].data_ptr(), 0, n(t),
                                  nullptr, B, A, stream), "rotate");
    });
    const double ce = timed([&](std::vector<torch::Tensor>& t) {
        check_rc(tpl_xform_center(dt, t[4].data_ptr(), t[0].data_ptr(), n(t), nullptr, B, A, stream),
                 "center");
    });
    const double cb = timed([&](std::vector<torch::Tensor>& t) {
        check_rc(tpl_xform_center_backward(dt, t[1].data_ptr(), t[4].data_ptr(), n(t), nullptr, B,
                                           A, stream), "center_backward");
    });
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return {tr, ro, ce, cb};
}

}  // namespace

// Angles2Coords entry points (a2c_ext.cpp), in the same module as in the
// reference (Layers/FullAtomModel/main.cpp:10-14).
void register_a2c(pybind11::module& m);
// PDB2CoordsUnordered (pdb_ext.cpp), likewise (main.cpp:21).
void register_pdb(pybind11::module& m);

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.doc() = "Coordinate transforms for MI355X: tensor glue over the tpl_xform C ABI";
    namespace py = pybind11;
    auto st = py::arg("status") = py::none();
    m.def("CoordsTranslateGPU_forward", &translate_forward, "Translate coordinates on GPU",
          py::arg("input_coords"), py::arg("output_coords"), py::arg("T"), py::arg("num_atoms"), st);
    m.def("CoordsTranslateGPU_backward", &translate_backward, "Backward of translate coordinates",
          py::arg("grad_output_coords"), py::arg("grad_input_coords"), py::arg("T"),
          py::arg("num_atoms"), st);
    m.def("CoordsRotateGPU_forward", &rotate_forward, "Rotate coordinates on GPU",
          py::arg("input_coords"), py::arg("output_coords"), py::arg("R"), py::arg("num_atoms"), st);
    m.def("CoordsRotateGPU_backward", &rotate_backward, "Backward of rotate coordinates on GPU",
          py::arg("grad_output_coords"), py::arg("grad_input_coords"), py::arg("R"),
          py::arg("num_atoms"), st);
    m.def("Coords2CenterGPU_forward", &center_forward, "Get coordinates center on GPU",
          py::arg("input_coords"), py::arg("output_center"), py::arg("num_atoms"), st);
    m.def("Coords2CenterGPU_backward", &center_backward, "Backward of coordinates center on GPU",
          py::arg("grad_output_center"), py::arg("grad_input_coords"), py::arg("num_atoms"), st);
    m.def("bench", &bench, "average event ms of translate, rotate, centre, centre-backward");
    register_a2c(m);
    register_pdb(m);
}
