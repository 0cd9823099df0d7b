// a2c_ext.cpp -- the Angles2Coords part of the pybind11 module
// `_FullAtomModel`: tensor glue over the C ABI in include/tpl_a2c.h, with the
// reference's entry-point names and argument orders
// (Layers/FullAtomModel/main.cpp:10-14,
// Layers/FullAtomModel/Angles2Coords/angles2coords_interface.cpp:8-144).
// The reference takes CPU tensors; these take GPU tensors (the sequences as a
// uint8 [batch, length] tensor on the same device).  Work goes to torch's
// current HIP stream of the angles' device.
#include <torch/extension.h>
#include <c10/hip/HIPStream.h>

#include <tuple>
#include <vector>

#include "tpl_a2c.h"

namespace {

#define A2C_CHECK_GPU(x)                                                      \
    TORCH_CHECK((x).is_cuda(), #x " must be a GPU tensor");                   \
    TORCH_CHECK((x).is_contiguous(), #x " must be contiguous")

void check_rc(int rc, const char* what) {
    TORCH_CHECK(rc == TPL_A2C_OK, what, " failed: ", tpl_a2c_strerror(rc));
}

int dtype_code(const torch::Tensor& t) {
    if (t.scalar_type() == torch::kFloat) return TPL_A2C_F32;
    if (t.scalar_type() == torch::kDouble) return TPL_A2C_F64;
    TORCH_CHECK(false, "angles must be float32 or float64, got ", t.scalar_type());
    return -1;
}

void check_common(const torch::Tensor& seqs, const torch::Tensor& angles) {
    A2C_CHECK_GPU(seqs);
    A2C_CHECK_GPU(angles);
    TORCH_CHECK(seqs.scalar_type() == torch::kByte && seqs.dim() == 2, "sequences must be uint8 [batch, length]");
    TORCH_CHECK(angles.dim() == 3 && angles.size(1) == 8, "angles must be [batch, 8, length], got ", angles.sizes());
    TORCH_CHECK(seqs.size(0) == angles.size(0), "sequences and angles differ in batch");
    TORCH_CHECK(seqs.device() == angles.device(), "sequences must be on the device of the angles");
    TORCH_CHECK(angles.size(0) <= INT32_MAX && angles.size(2) <= INT32_MAX / 8 && seqs.size(1) <= INT32_MAX,
                "inputs too large");
}

uint32_t* status_ptr(const c10::optional<torch::Tensor>& status, const torch::Tensor& like) {
    if (!status.has_value() || !status->defined()) return nullptr;
    A2C_CHECK_GPU(*status);
    TORCH_CHECK(status->scalar_type() == torch::kInt && status->numel() >= 1, "status must be an int32 tensor");
    TORCH_CHECK(status->device() == like.device(), "status must be on the device of the angles");
    return reinterpret_cast<uint32_t*>(status->data_ptr());
}

// The optional scheduling order (tpl_a2c_forward_ordered): int32 [batch] on
// the angles' device.
const int32_t* order_ptr(const c10::optional<torch::Tensor>& order, const torch::Tensor& like) {
    if (!order.has_value() || !order->defined()) return nullptr;
    A2C_CHECK_GPU(*order);
    TORCH_CHECK(order->scalar_type() == torch::kInt && order->dim() == 1 && order->size(0) == like.size(0),
                "order must be int32 [batch]");
    TORCH_CHECK(order->device() == like.device(), "order must be on the device of the angles");
    return order->data_ptr<int32_t>();
}

hipStream_t stream_of(const torch::Tensor& t) {
    return c10::hip::getCurrentHIPStream(t.device().index()).stream();
}

void forward(torch::Tensor sequences, torch::Tensor input_angles, torch::Tensor output_coords,
             torch::Tensor res_names, torch::Tensor atom_names, c10::optional<torch::Tensor> status,
             c10::optional<torch::Tensor> order) {
    check_common(sequences, input_angles);
    A2C_CHECK_GPU(output_coords);
    A2C_CHECK_GPU(res_names);
    A2C_CHECK_GPU(atom_names);
    const int64_t B = input_angles.size(0);
    TORCH_CHECK(output_coords.scalar_type() == input_angles.scalar_type(), "coordinates must have the dtype of the angles");
    TORCH_CHECK(output_coords.dim() == 2 && output_coords.size(0) == B && output_coords.size(1) % 3 == 0,
                "output_coords must be [batch, 3*atoms]");
    const int64_t A = output_coords.size(1) / 3;
    TORCH_CHECK(A <= INT32_MAX / 3, "output_coords too large");
    for (const auto* n : {&res_names, &atom_names}) {
        TORCH_CHECK(n->scalar_type() == torch::kByte && n->dim() == 3 && n->size(0) == B && n->size(1) == A &&
                        n->size(2) == 4,
                    "names must be uint8 [batch, atoms, 4]");
        TORCH_CHECK(n->device() == input_angles.device(), "names must be on the device of the angles");
    }
    TORCH_CHECK(output_coords.device() == input_angles.device(), "coordinates must be on the device of the angles");
    c10::DeviceGuard guard(input_angles.device());
    check_rc(tpl_a2c_forward_ordered(order_ptr(order, input_angles), dtype_code(input_angles),
                                     sequences.data_ptr<uint8_t>(), (int)sequences.size(1), input_angles.data_ptr(),
                                     (int)input_angles.size(2), output_coords.data_ptr(),
                                     res_names.data_ptr<uint8_t>(), atom_names.data_ptr<uint8_t>(), (int)A,
                                     status_ptr(status, input_angles), (int)B, stream_of(input_angles)),
             "Angles2Coords_forward");
}

// From the forward's coordinates (what the Python layer saves).
void backward_coords(torch::Tensor grad_atoms, torch::Tensor grad_angles, torch::Tensor sequences,
                     torch::Tensor coords, c10::optional<torch::Tensor> status,
                     c10::optional<torch::Tensor> order) {
    check_common(sequences, grad_angles);
    A2C_CHECK_GPU(grad_atoms);
    A2C_CHECK_GPU(coords);
    const int64_t B = grad_angles.size(0);
    TORCH_CHECK(grad_atoms.scalar_type() == grad_angles.scalar_type() &&
                    coords.scalar_type() == grad_angles.scalar_type(),
                "gradients and coordinates must have one dtype");
    TORCH_CHECK(grad_atoms.dim() == 2 && grad_atoms.size(0) == B && grad_atoms.size(1) % 3 == 0,
                "grad_atoms must be [batch, 3*atoms]");
    TORCH_CHECK(coords.sizes() == grad_atoms.sizes(), "coords must have the shape of grad_atoms");
    TORCH_CHECK(grad_atoms.device() == grad_angles.device() && coords.device() == grad_angles.device(),
                "gradients and coordinates must be on one device");
    const int64_t A = grad_atoms.size(1) / 3;
    TORCH_CHECK(A <= INT32_MAX / 3, "grad_atoms too large");
    c10::DeviceGuard guard(grad_angles.device());
    check_rc(tpl_a2c_backward_form(TPL_A2C_BWD_AUTO, order_ptr(order, grad_angles), dtype_code(grad_angles),
                                   sequences.data_ptr<uint8_t>(), (int)sequences.size(1), coords.data_ptr(),
                                   grad_atoms.data_ptr(), (int)A, grad_angles.data_ptr(), (int)grad_angles.size(2),
                                   status_ptr(status, grad_angles), (int)B, stream_of(grad_angles)),
             "Angles2Coords_backward");
}

// The reference's signature (angles, not coordinates): the coordinates are
// computed again first, as the reference rebuilds its conformation.
void backward(torch::Tensor grad_atoms, torch::Tensor grad_angles, torch::Tensor sequences,
              torch::Tensor input_angles, c10::optional<torch::Tensor> status) {
    check_common(sequences, input_angles);
    A2C_CHECK_GPU(grad_atoms);
    TORCH_CHECK(grad_angles.sizes() == input_angles.sizes(), "grad_angles must have the shape of the angles");
    TORCH_CHECK(grad_atoms.dim() == 2 && grad_atoms.size(0) == input_angles.size(0) && grad_atoms.size(1) % 3 == 0,
                "grad_atoms must be [batch, 3*atoms]");
    const int64_t A = grad_atoms.size(1) / 3;
    TORCH_CHECK(A <= INT32_MAX / 3, "grad_atoms too large");
    c10::DeviceGuard guard(input_angles.device());
    auto coords = torch::empty_like(grad_atoms);
    check_rc(tpl_a2c_forward(dtype_code(input_angles), sequences.data_ptr<uint8_t>(), (int)sequences.size(1),
                             input_angles.data_ptr(), (int)input_angles.size(2), coords.data_ptr(), nullptr,
                             nullptr, (int)A, status_ptr(status, input_angles), (int)input_angles.size(0),
                             stream_of(input_angles)),
             "Angles2Coords_backward");
    backward_coords(grad_atoms, grad_angles, sequences, coords, status, c10::nullopt);
}

int num_atoms(const std::string& sequence) {
    const int n = tpl_a2c_num_atoms(sequence.c_str());
    TORCH_CHECK(n >= 0, "getSeqNumAtoms: sequence holds a letter outside the 20 amino acids: ", sequence);
    return n;
}

// Timing loop for bench.py: sets of [sequences, angles, coords, grad_coords,
// grad_angles] (and optionally the scheduling order, int32 [batch]) rotated
// per launch (no names, no status).  Returns the average
// HIP-event ms of (fwd+bwd step, fwd alone, bwd alone) over `steps` launches
// each, after `warmup`.
std::tuple<double, double, double> bench(std::vector<std::vector<torch::Tensor>> sets, int64_t steps,
                                         int64_t warmup) {
    TORCH_CHECK(!sets.empty() && steps > 0 && warmup >= 0, "bench: bad arguments");
    for (auto& t : sets) {
        TORCH_CHECK(t.size() == 5 || t.size() == 6, "bench: a set has 5 tensors (6 with an order)");
        if (t.size() == 6) (void)order_ptr(t[5], t[1]);
        check_common(t[0], t[1]);
        for (int k = 2; k < 5; ++k) {
            A2C_CHECK_GPU(t[k]);
            TORCH_CHECK(t[k].scalar_type() == t[1].scalar_type(), "bench: dtype mismatch");
        }
        TORCH_CHECK(t[2].sizes() == t[3].sizes() && t[4].sizes() == t[1].sizes(), "bench: shape mismatch");
    }
    c10::DeviceGuard guard(sets[0][1].device());
    hipStream_t stream = stream_of(sets[0][1]);
    const int dt = dtype_code(sets[0][1]);
    auto fwd = [&](std::vector<torch::Tensor>& t) {
        const int32_t* order = t.size() == 6 ? t[5].data_ptr<int32_t>() : nullptr;
        check_rc(tpl_a2c_forward_ordered(order, dt, t[0].data_ptr<uint8_t>(), (int)t[0].size(1), t[1].data_ptr(),
                                         (int)t[1].size(2), t[2].data_ptr(), nullptr, nullptr,
                                         (int)(t[2].size(1) / 3), nullptr, (int)t[1].size(0), stream),
                 "forward");
    };
    auto bwd = [&](std::vector<torch::Tensor>& t) {
        const int32_t* order = t.size() == 6 ? t[5].data_ptr<int32_t>() : nullptr;
        check_rc(tpl_a2c_backward_form(TPL_A2C_BWD_AUTO, order, dt, t[0].data_ptr<uint8_t>(), (int)t[0].size(1),
                                       t[2].data_ptr(), t[3].data_ptr(), (int)(t[3].size(1) / 3), t[4].data_ptr(),
                                       (int)t[4].size(2), nullptr, (int)t[1].size(0), stream),
                 "backward");
    };
    hipEvent_t e0, e1;
    TORCH_CHECK(hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess, "hipEventCreate failed");
    auto timed = [&](int which) {
        auto body = [&](std::vector<torch::Tensor>& t) {
            if (which != 2) fwd(t);
            if (which != 1) bwd(t);
        };
        for (int64_t i = 0; i < warmup; ++i) body(sets[i % sets.size()]);
        TORCH_CHECK(hipEventRecord(e0, stream) == hipSuccess, "hipEventRecord failed");
        for (int64_t i = 0; i < steps; ++i) body(sets[i % sets.size()]);
        TORCH_CHECK(hipEventRecord(e1, stream) == hipSuccess, "hipEventRecord failed");
        TORCH_CHECK(hipEventSynchronize(e1) == hipSuccess, "hipEventSynchronize failed");
        float ms = 0.f;
        TORCH_CHECK(hipEventElapsedTime(&ms, e0, e1) == hipSuccess, "hipEventElapsedTime failed");
        return (double)ms / steps;
    };
    const double step = timed(0), f = timed(1), b = timed(2);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return {step, f, b};
}

}  // namespace

void register_a2c(pybind11::module& m) {
    namespace py = pybind11;
    auto st = py::arg("status") = py::none();
    auto od = py::arg("order") = py::none();
    m.def("Angles2Coords_forward", &forward, "full-atom coordinates and names from angles (GPU)",
          py::arg("sequences"), py::arg("input_angles"), py::arg("output_coords"), py::arg("res_names"),
          py::arg("atom_names"), st, od);
    m.def("Angles2Coords_backward", &backward, "gradient of the coordinates with respect to the angles (GPU)",
          py::arg("grad_atoms"), py::arg("grad_angles"), py::arg("sequences"), py::arg("input_angles"), st);
    m.def("Angles2Coords_backward_coords", &backward_coords,
          "the same from the forward's coordinates (GPU)", py::arg("grad_atoms"), py::arg("grad_angles"),
          py::arg("sequences"), py::arg("coords"), st, od);
    m.def("getSeqNumAtoms", &num_atoms, "atoms of a sequence", py::arg("sequence"));
    m.def("a2c_bench", &bench, "average event ms of Angles2Coords (step, fwd, bwd)");
}
