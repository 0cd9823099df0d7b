// a2b_ext.cpp -- pybind11 module `_ReducedModel`: the tensor-facing glue over
// the C ABI in include/tpl_a2b.h.
//
// Mirrors the reference module of the same name (Layers/ReducedModel/main.cpp:4-9)
// and its checks (Layers/ReducedModel/Angles2Backbone/angles2backbone_interface.cpp:12-18,
// Math/nUtil.h:49-57): device, contiguity and dtype checks raise RuntimeError.
// The ndim check that the reference's no-op ERROR() macro skips is enforced.
// Work goes to torch's current HIP stream of the input's device, not the
// legacy default stream (cBackboneProteinCUDAKernels.cu:166,172,177).
#include <torch/extension.h>
#include <c10/hip/HIPStream.h>

#include <chrono>
#include <string>
#include <tuple>
#include <vector>

#include "tpl_a2b.h"
#include "tpl_a2b_cpu.h"

namespace {

#define A2B_CHECK_GPU(x)                                                      \
    TORCH_CHECK((x).is_cuda(), #x " must be a GPU tensor");                   \
    TORCH_CHECK((x).is_contiguous(), #x " must be contiguous")

void check_rc(int rc, const char* what) {
    TORCH_CHECK(rc == TPL_A2B_OK, what, " failed: ", tpl_a2b_strerror(rc));
}

int dtype_code(const torch::Tensor& t) {
    if (t.scalar_type() == torch::kFloat) return TPL_A2B_F32;
    if (t.scalar_type() == torch::kDouble) return TPL_A2B_F64;
    TORCH_CHECK(false, "Angles2Backbone: angles must be float32 or float64, got ",
                t.scalar_type());
    return -1;
}

tpl_a2b_config to_config(const std::vector<int64_t>& c) {
    TORCH_CHECK(c.size() == 2, "config must be (fwd_res_per_lane, bwd_res_per_lane)");
    tpl_a2b_config cfg{};
    cfg.fwd_res_per_lane = (int32_t)c[0];
    cfg.bwd_res_per_lane = (int32_t)c[1];
    cfg.flags = 0;
    cfg.reserved = 0;
    TORCH_CHECK(tpl_a2b_chunk_residues(&cfg, 0) > 0, "unsupported config (fwd_res_per_lane=", c[0],
                ", bwd_res_per_lane=", c[1], ")");
    return cfg;
}

std::vector<int64_t> from_config(const tpl_a2b_config& c) {
    return {c.fwd_res_per_lane, c.bwd_res_per_lane};
}

// Validates angles [B,3,L] + lengths int32 [B] on one device; returns (B, L).
std::pair<int, int> check_inputs(const torch::Tensor& angles, const torch::Tensor& lengths) {
    A2B_CHECK_GPU(angles);
    A2B_CHECK_GPU(lengths);
    TORCH_CHECK(lengths.scalar_type() == torch::kInt, "angles_length must be int32");
    TORCH_CHECK(angles.dim() == 3 && angles.size(1) == 3,
                "input_angles must be [batch, 3, max_length], got ", angles.sizes());
    TORCH_CHECK(lengths.dim() == 1 && lengths.size(0) == angles.size(0),
                "angles_length must be [batch]");
    TORCH_CHECK(lengths.device() == angles.device(), "angles_length must be on the device of input_angles");
    TORCH_CHECK(angles.size(0) <= INT32_MAX && angles.size(2) <= INT32_MAX, "input too large");
    return {(int)angles.size(0), (int)angles.size(2)};
}

// The host path: CPU tensors, contiguous, float64 only (the reference's CPU
// Function is float64-only, Angles2Backbone.py:64-65), lengths int32.
std::pair<int, int> check_cpu_inputs(const torch::Tensor& angles, const torch::Tensor& lengths) {
    TORCH_CHECK(!angles.is_cuda() && angles.is_contiguous(), "input_angles must be a contiguous CPU tensor");
    TORCH_CHECK(angles.scalar_type() == torch::kDouble,
                "Angles2Backbone on the CPU computes in float64: input must be float64, got ",
                angles.scalar_type());
    TORCH_CHECK(!lengths.is_cuda() && lengths.is_contiguous(), "angles_length must be a contiguous CPU tensor");
    TORCH_CHECK(lengths.scalar_type() == torch::kInt, "angles_length must be int32");
    TORCH_CHECK(angles.dim() == 3 && angles.size(1) == 3,
                "input_angles must be [batch, 3, max_length], got ", angles.sizes());
    TORCH_CHECK(lengths.dim() == 1 && lengths.size(0) == angles.size(0), "angles_length must be [batch]");
    TORCH_CHECK(angles.size(0) <= INT32_MAX && angles.size(2) <= INT32_MAX, "input too large");
    return {(int)angles.size(0), (int)angles.size(2)};
}

void check_cpu_out(const torch::Tensor& t, std::vector<int64_t> shape, const char* name) {
    TORCH_CHECK(!t.is_cuda() && t.is_contiguous(), name, " must be a contiguous CPU tensor");
    TORCH_CHECK(t.scalar_type() == torch::kDouble, name, " must be float64");
    TORCH_CHECK(t.sizes() == torch::IntArrayRef(shape), name, " has shape ", t.sizes());
}

void check_out(const torch::Tensor& t, const torch::Tensor& like, std::vector<int64_t> shape,
               const char* name) {
    A2B_CHECK_GPU(t);
    TORCH_CHECK(t.scalar_type() == like.scalar_type(), name, " must have the dtype of the angles");
    TORCH_CHECK(t.device() == like.device(), name, " must be on the device of the angles");
    TORCH_CHECK(t.sizes() == torch::IntArrayRef(shape), name, " has shape ", t.sizes());
}

void* opt_ptr(const torch::Tensor& t, size_t need, const char* name, const torch::Tensor& like) {
    if (need == 0) return nullptr;
    A2B_CHECK_GPU(t);
    TORCH_CHECK(t.device() == like.device(), name, " must be on the device of the angles");
    TORCH_CHECK((size_t)t.nbytes() >= need, name, " needs ", need, " bytes, has ", t.nbytes());
    return t.data_ptr();
}

uint32_t* status_ptr(const c10::optional<torch::Tensor>& status, const torch::Tensor& like) {
    if (!status.has_value() || !status->defined()) return nullptr;
    A2B_CHECK_GPU(*status);
    TORCH_CHECK(status->scalar_type() == torch::kInt && status->numel() >= 1,
                "status must be an int32 tensor");
    TORCH_CHECK(status->device() == like.device(), "status must be on the device of the angles");
    return reinterpret_cast<uint32_t*>(status->data_ptr());
}

hipStream_t stream_of(const torch::Tensor& t) {
    return c10::hip::getCurrentHIPStream(t.device().index()).stream();
}

// ----------------------------------------------------------------- bindings

// Angles2BackboneGPU_forward (reference angles2backbone_interface.cpp:7-26):
// (input_angles, output_coords, angles_length, A) with A replaced by the
// forward->backward checkpoints (None: no backward follows), plus the
// look-back words (zero before the first call; left zero), optional status
// and config.
void forward(torch::Tensor input_angles, torch::Tensor output_coords,
             torch::Tensor angles_length, c10::optional<torch::Tensor> state, torch::Tensor sync,
             c10::optional<torch::Tensor> status, std::vector<int64_t> config) {
    auto [B, L] = check_inputs(input_angles, angles_length);
    check_out(output_coords, input_angles, {B, 9 * (int64_t)L}, "output_coords");
    const tpl_a2b_config cfg = to_config(config);
    const int dt = dtype_code(input_angles);
    c10::DeviceGuard guard(input_angles.device());
    void* st = nullptr;
    if (state.has_value() && state->defined())
        st = opt_ptr(*state, tpl_a2b_state_bytes(dt, B, L, &cfg), "state", input_angles);
    void* sy = opt_ptr(sync, tpl_a2b_sync_bytes(dt, B, L, &cfg), "sync", input_angles);
    check_rc(tpl_a2b_forward(dt, input_angles.data_ptr(), output_coords.data_ptr(),
                             angles_length.data_ptr<int32_t>(), st, sy,
                             status_ptr(status, input_angles), B, L, &cfg,
                             stream_of(input_angles)),
             "Angles2BackboneGPU_forward");
}

// Angles2BackboneGPU_backward (reference angles2backbone_interface.cpp:27-56):
// (gradInput, gradOutput, input_angles, angles_length, A, dr_dangle) with A ->
// the forward's checkpoints and dr_dangle -> the look-back words.
void backward(torch::Tensor grad_input, torch::Tensor grad_output, torch::Tensor input_angles,
              torch::Tensor angles_length, torch::Tensor state, torch::Tensor sync,
              c10::optional<torch::Tensor> status, std::vector<int64_t> config) {
    auto [B, L] = check_inputs(input_angles, angles_length);
    check_out(grad_input, input_angles, {B, 3, L}, "gradInput");
    check_out(grad_output, input_angles, {B, 9 * (int64_t)L}, "gradOutput");
    const tpl_a2b_config cfg = to_config(config);
    const int dt = dtype_code(input_angles);
    c10::DeviceGuard guard(input_angles.device());
    void* st = opt_ptr(state, tpl_a2b_state_bytes(dt, B, L, &cfg), "state", input_angles);
    void* sy = opt_ptr(sync, tpl_a2b_sync_bytes(dt, B, L, &cfg), "sync", input_angles);
    check_rc(tpl_a2b_backward(dt, grad_input.data_ptr(), grad_output.data_ptr(),
                              input_angles.data_ptr(), angles_length.data_ptr<int32_t>(), st, sy,
                              status_ptr(status, input_angles), B, L, &cfg,
                              stream_of(input_angles)),
             "Angles2BackboneGPU_backward");
}

void sync_clear(torch::Tensor sync) {
    A2B_CHECK_GPU(sync);
    c10::DeviceGuard guard(sync.device());
    check_rc(tpl_a2b_sync_clear(sync.data_ptr(), (size_t)sync.nbytes() & ~size_t(3), stream_of(sync)),
             "sync_clear");
}

// Angles2BackboneCPU_forward (reference angles2backbone_interface.cpp:60-80):
// (input_angles, output_coords, angles_length) on the host, float64; the
// reference's frame buffer A is not needed.  Returns the status bits.
int64_t cpu_forward(torch::Tensor input_angles, torch::Tensor output_coords,
                    torch::Tensor angles_length, int64_t threads) {
    auto [B, L] = check_cpu_inputs(input_angles, angles_length);
    check_cpu_out(output_coords, {B, 9 * (int64_t)L}, "output_coords");
    uint32_t status = 0;
    check_rc(tpl_a2b_cpu_forward(input_angles.data_ptr<double>(), output_coords.data_ptr<double>(),
                                 angles_length.data_ptr<int32_t>(), &status, B, L, (int)threads),
             "Angles2BackboneCPU_forward");
    return status;
}

// Angles2BackboneCPU_backward (reference angles2backbone_interface.cpp:83-111):
// (gradInput, gradOutput, output_coords, angles_length): the forward's
// coordinates stand where input_angles and A stood; no dr_dangle.  Returns
// the status bits.
int64_t cpu_backward(torch::Tensor grad_input, torch::Tensor grad_output,
                     torch::Tensor output_coords, torch::Tensor angles_length, int64_t threads) {
    TORCH_CHECK(!grad_input.is_cuda(), "gradInput must be a CPU tensor");
    TORCH_CHECK(grad_input.dim() == 3, "gradInput must be [batch, 3, max_length]");
    auto [B, L] = check_cpu_inputs(grad_input, angles_length);
    check_cpu_out(grad_output, {B, 9 * (int64_t)L}, "gradOutput");
    check_cpu_out(output_coords, {B, 9 * (int64_t)L}, "output_coords");
    uint32_t status = 0;
    check_rc(tpl_a2b_cpu_backward(grad_input.data_ptr<double>(), grad_output.data_ptr<double>(),
                                  output_coords.data_ptr<double>(),
                                  angles_length.data_ptr<int32_t>(), &status, B, L, (int)threads),
             "Angles2BackboneCPU_backward");
    return status;
}

void forward_plain(torch::Tensor input_angles, torch::Tensor output_coords,
                   torch::Tensor angles_length, torch::Tensor frames,
                   c10::optional<torch::Tensor> status) {
    auto [B, L] = check_inputs(input_angles, angles_length);
    check_out(output_coords, input_angles, {B, 9 * (int64_t)L}, "output_coords");
    const int dt = dtype_code(input_angles);
    c10::DeviceGuard guard(input_angles.device());
    void* fr = opt_ptr(frames, tpl_a2b_plain_frames_bytes(dt, B, L), "frames", input_angles);
    check_rc(tpl_a2b_forward_plain(dt, input_angles.data_ptr(), output_coords.data_ptr(),
                                   angles_length.data_ptr<int32_t>(), fr,
                                   status_ptr(status, input_angles), B, L,
                                   stream_of(input_angles)),
             "Angles2BackboneGPU_forward_plain");
}

void backward_plain(torch::Tensor grad_input, torch::Tensor grad_output,
                    torch::Tensor input_angles, torch::Tensor angles_length,
                    torch::Tensor frames, c10::optional<torch::Tensor> status) {
    auto [B, L] = check_inputs(input_angles, angles_length);
    check_out(grad_input, input_angles, {B, 3, L}, "gradInput");
    check_out(grad_output, input_angles, {B, 9 * (int64_t)L}, "gradOutput");
    const int dt = dtype_code(input_angles);
    c10::DeviceGuard guard(input_angles.device());
    void* fr = opt_ptr(frames, tpl_a2b_plain_frames_bytes(dt, B, L), "frames", input_angles);
    check_rc(tpl_a2b_backward_plain(dt, grad_input.data_ptr(), grad_output.data_ptr(),
                                    input_angles.data_ptr(), angles_length.data_ptr<int32_t>(),
                                    fr, status_ptr(status, input_angles), B, L,
                                    stream_of(input_angles)),
             "Angles2BackboneGPU_backward_plain");
}

std::tuple<int64_t, int64_t> workspace_bytes(bool f64, int64_t B, int64_t L,
                                             std::vector<int64_t> config) {
    const tpl_a2b_config cfg = to_config(config);
    const int dt = f64 ? TPL_A2B_F64 : TPL_A2B_F32;
    return {(int64_t)tpl_a2b_state_bytes(dt, (int)B, (int)L, &cfg),
            (int64_t)tpl_a2b_sync_bytes(dt, (int)B, (int)L, &cfg)};
}

int64_t plain_frames_bytes(bool f64, int64_t B, int64_t L) {
    return (int64_t)tpl_a2b_plain_frames_bytes(f64 ? TPL_A2B_F64 : TPL_A2B_F32, (int)B, (int)L);
}

std::vector<int64_t> default_config(int64_t B, int64_t L) {
    tpl_a2b_config cfg{};
    check_rc(tpl_a2b_default_config((int)B, (int)L, &cfg), "default_config");
    return from_config(cfg);
}

std::tuple<std::vector<int64_t>, std::vector<double>> tune(
    torch::Tensor input_angles, torch::Tensor angles_length, torch::Tensor grad_output,
    torch::Tensor output_coords, torch::Tensor grad_input, torch::Tensor state, torch::Tensor sync,
    std::vector<std::vector<int64_t>> candidates, int64_t reps) {
    auto [B, L] = check_inputs(input_angles, angles_length);
    check_out(output_coords, input_angles, {B, 9 * (int64_t)L}, "output_coords");
    check_out(grad_input, input_angles, {B, 3, L}, "gradInput");
    check_out(grad_output, input_angles, {B, 9 * (int64_t)L}, "gradOutput");
    TORCH_CHECK(!candidates.empty(), "no candidates");
    std::vector<tpl_a2b_config> cands;
    size_t need_state = 0, need_sync = 0;
    for (auto& c : candidates) {
        cands.push_back(to_config(c));
        need_state = std::max(need_state, tpl_a2b_state_bytes(dtype_code(input_angles), B, L, &cands.back()));
        need_sync = std::max(need_sync, tpl_a2b_sync_bytes(dtype_code(input_angles), B, L, &cands.back()));
    }
    c10::DeviceGuard guard(input_angles.device());
    void* st = opt_ptr(state, need_state, "state", input_angles);
    void* sy = opt_ptr(sync, need_sync, "sync", input_angles);
    std::vector<float> ms(cands.size(), -1.f);
    tpl_a2b_config best{};
    check_rc(tpl_a2b_tune(dtype_code(input_angles), input_angles.data_ptr(),
                          angles_length.data_ptr<int32_t>(), grad_output.data_ptr(),
                          output_coords.data_ptr(), grad_input.data_ptr(), st, sy, B, L,
                          cands.data(), (int)cands.size(), (int)reps, &best, ms.data(),
                          stream_of(input_angles)),
             "tune");
    return {from_config(best), std::vector<double>(ms.begin(), ms.end())};
}

// Times `steps` back-to-back steps (forward then backward launch) on the
// current stream, rotating over the buffer sets, after `warmup` untimed steps.
// Each set is [angles, lengths, grad_coords, coords, grad_angles, state, sync]
// (or, for plain=true, [..., frames, unused]); sync must be zero before the
// first call (the calls leave it zero).  Returns (event ms, wall ms).
std::tuple<double, double> bench_steps(std::vector<std::vector<torch::Tensor>> sets,
                                       std::vector<int64_t> config, int64_t steps,
                                       int64_t warmup, bool plain) {
    TORCH_CHECK(!sets.empty() && steps > 0 && warmup >= 0, "bench_steps: bad arguments");
    for (auto& s : sets) TORCH_CHECK(s.size() == 7, "bench_steps: a set has 7 tensors");
    const tpl_a2b_config cfg = to_config(config);
    const torch::Tensor& a0 = sets[0][0];
    auto [B, L] = check_inputs(a0, sets[0][1]);
    const int dt = dtype_code(a0);
    c10::DeviceGuard guard(a0.device());
    hipStream_t stream = stream_of(a0);
    auto one = [&](const std::vector<torch::Tensor>& s) {
        int rc;
        if (plain) {
            rc = tpl_a2b_forward_plain(dt, s[0].data_ptr(), s[3].data_ptr(), s[1].data_ptr<int32_t>(),
                                       s[5].data_ptr(), nullptr, B, L, stream);
            if (rc == TPL_A2B_OK)
                rc = tpl_a2b_backward_plain(dt, s[4].data_ptr(), s[2].data_ptr(), s[0].data_ptr(),
                                            s[1].data_ptr<int32_t>(), s[5].data_ptr(), nullptr, B,
                                            L, stream);
        } else {
            rc = tpl_a2b_forward(dt, s[0].data_ptr(), s[3].data_ptr(), s[1].data_ptr<int32_t>(),
                                 s[5].data_ptr(), s[6].data_ptr(), nullptr, B, L, &cfg, stream);
            if (rc == TPL_A2B_OK)
                rc = tpl_a2b_backward(dt, s[4].data_ptr(), s[2].data_ptr(), s[0].data_ptr(),
                                      s[1].data_ptr<int32_t>(), s[5].data_ptr(), s[6].data_ptr(),
                                      nullptr, B, L, &cfg, stream);
        }
        check_rc(rc, "bench_steps");
    };
    for (auto& s : sets) {  // every set is checked once, as the bindings above do
        check_inputs(s[0], s[1]);
        check_out(s[2], a0, {B, 9 * (int64_t)L}, "gradOutput");
        check_out(s[3], a0, {B, 9 * (int64_t)L}, "output_coords");
        check_out(s[4], a0, {B, 3, L}, "gradInput");
        if (plain) {
            opt_ptr(s[5], tpl_a2b_plain_frames_bytes(dt, B, L), "frames", a0);
        } else {
            opt_ptr(s[5], tpl_a2b_state_bytes(dt, B, L, &cfg), "state", a0);
            opt_ptr(s[6], tpl_a2b_sync_bytes(dt, B, L, &cfg), "sync", a0);
        }
    }
    for (int64_t i = 0; i < warmup; ++i) one(sets[i % sets.size()]);
    hipEvent_t e0, e1;
    TORCH_CHECK(hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess,
                "hipEventCreate failed");
    TORCH_CHECK(hipStreamSynchronize(stream) == hipSuccess, "hipStreamSynchronize failed");
    const auto t0 = std::chrono::steady_clock::now();
    TORCH_CHECK(hipEventRecord(e0, stream) == hipSuccess, "hipEventRecord failed");
    for (int64_t i = 0; i < steps; ++i) one(sets[i % sets.size()]);
    TORCH_CHECK(hipEventRecord(e1, stream) == hipSuccess, "hipEventRecord failed");
    TORCH_CHECK(hipEventSynchronize(e1) == hipSuccess, "hipEventSynchronize failed");
    const auto t1 = std::chrono::steady_clock::now();
    float ms = 0.f;
    TORCH_CHECK(hipEventElapsedTime(&ms, e0, e1) == hipSuccess, "hipEventElapsedTime failed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return {(double)ms, std::chrono::duration<double, std::milli>(t1 - t0).count()};
}

// The same step captured as a HIP graph: for each buffer set, one forward and
// one backward launch are captured on a private stream (thread-local capture
// mode) and instantiated; the graphs are then replayed back to back on the
// current stream, `warmup` untimed and `steps` timed.  Returns (event ms,
// wall ms).  Nothing is allocated or cleared inside the captured work.
// One HIP graph holds a forward + backward per buffer set, in order; it is
// replayed ceil(steps / sets) times (graph launches amortised over the sets).
// Returns (event ms, wall ms, steps replayed).
std::tuple<double, double, int64_t> bench_graph(std::vector<std::vector<torch::Tensor>> sets,
                                                std::vector<int64_t> config, int64_t steps,
                                                int64_t warmup) {
    TORCH_CHECK(!sets.empty() && steps >= 0 && warmup >= 0, "bench_graph: bad arguments");
    for (auto& s : sets) TORCH_CHECK(s.size() == 7, "bench_graph: a set has 7 tensors");
    const tpl_a2b_config cfg = to_config(config);
    const torch::Tensor& a0 = sets[0][0];
    auto [B, L] = check_inputs(a0, sets[0][1]);
    const int dt = dtype_code(a0);
    c10::DeviceGuard guard(a0.device());
    for (auto& s : sets) {
        check_inputs(s[0], s[1]);
        check_out(s[2], a0, {B, 9 * (int64_t)L}, "gradOutput");
        check_out(s[3], a0, {B, 9 * (int64_t)L}, "output_coords");
        check_out(s[4], a0, {B, 3, L}, "gradInput");
        opt_ptr(s[5], tpl_a2b_state_bytes(dt, B, L, &cfg), "state", a0);
        opt_ptr(s[6], tpl_a2b_sync_bytes(dt, B, L, &cfg), "sync", a0);
    }
    hipStream_t stream = stream_of(a0);
    hipStream_t cs;
    TORCH_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) == hipSuccess,
                "hipStreamCreateWithFlags failed");
    std::vector<hipGraphExec_t> execs;
    std::string err;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        err = "hipStreamBeginCapture failed";
    } else {
        int rc = TPL_A2B_OK;
        for (auto& s : sets) {
            rc = tpl_a2b_forward(dt, s[0].data_ptr(), s[3].data_ptr(), s[1].data_ptr<int32_t>(),
                                 s[5].data_ptr(), s[6].data_ptr(), nullptr, B, L, &cfg, cs);
            if (rc == TPL_A2B_OK)
                rc = tpl_a2b_backward(dt, s[4].data_ptr(), s[2].data_ptr(), s[0].data_ptr(),
                                      s[1].data_ptr<int32_t>(), s[5].data_ptr(), s[6].data_ptr(),
                                      nullptr, B, L, &cfg, cs);
            if (rc != TPL_A2B_OK) break;
        }
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(cs, &g);
        if (rc != TPL_A2B_OK || ec != hipSuccess || !g) {
            err = rc != TPL_A2B_OK ? tpl_a2b_strerror(rc) : "hipStreamEndCapture failed";
        } else {
            hipGraphExec_t e = nullptr;
            if (hipGraphInstantiate(&e, g, nullptr, nullptr, 0) != hipSuccess)
                err = "hipGraphInstantiate failed";
            else
                execs.push_back(e);
        }
        if (g) (void)hipGraphDestroy(g);
    }
    auto cleanup = [&]() {
        for (auto e : execs) (void)hipGraphExecDestroy(e);
        (void)hipStreamDestroy(cs);
    };
    if (!err.empty()) {
        cleanup();
        TORCH_CHECK(false, "bench_graph: ", err);
    }
    const int64_t per = (int64_t)sets.size();
    const int64_t reps = (steps + per - 1) / per, wreps = (warmup + per - 1) / per;
    auto replay = [&]() {
        TORCH_CHECK(hipGraphLaunch(execs[0], stream) == hipSuccess, "hipGraphLaunch failed");
    };
    for (int64_t i = 0; i < wreps; ++i) replay();
    hipEvent_t e0, e1;
    TORCH_CHECK(hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess,
                "hipEventCreate failed");
    TORCH_CHECK(hipStreamSynchronize(stream) == hipSuccess, "hipStreamSynchronize failed");
    const auto t0 = std::chrono::steady_clock::now();
    TORCH_CHECK(hipEventRecord(e0, stream) == hipSuccess, "hipEventRecord failed");
    for (int64_t i = 0; i < reps; ++i) replay();
    TORCH_CHECK(hipEventRecord(e1, stream) == hipSuccess, "hipEventRecord failed");
    TORCH_CHECK(hipEventSynchronize(e1) == hipSuccess, "hipEventSynchronize failed");
    const auto t1 = std::chrono::steady_clock::now();
    float ms = 0.f;
    TORCH_CHECK(hipEventElapsedTime(&ms, e0, e1) == hipSuccess, "hipEventElapsedTime failed");
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cleanup();
    return {(double)ms, std::chrono::duration<double, std::milli>(t1 - t0).count(), reps * per};
}

// One forward+backward step over fixed buffers, captured once into a HIP
// graph (on a private stream, thread-local capture mode) and replayed on the
// current stream.  Each replay reads angles, lengths and grad_output from
// their memory and writes output_coords and grad_input; the look-back words
// (sync) must be zero before the first replay, and every replay leaves them
// zero.  The tensors are held for the graph's lifetime.
class StepGraph {
  public:
    StepGraph(torch::Tensor angles, torch::Tensor lengths, torch::Tensor grad_output,
              torch::Tensor output_coords, torch::Tensor grad_input, torch::Tensor state,
              torch::Tensor sync, std::vector<int64_t> config)
        : held_{angles, lengths, grad_output, output_coords, grad_input, state, sync} {
        const tpl_a2b_config cfg = to_config(config);
        auto [B, L] = check_inputs(angles, lengths);
        const int dt = dtype_code(angles);
        c10::DeviceGuard guard(angles.device());
        device_ = angles.device().index();
        check_out(grad_output, angles, {B, 9 * (int64_t)L}, "gradOutput");
        check_out(output_coords, angles, {B, 9 * (int64_t)L}, "output_coords");
        check_out(grad_input, angles, {B, 3, L}, "gradInput");
        void* st = opt_ptr(state, tpl_a2b_state_bytes(dt, B, L, &cfg), "state", angles);
        void* sy = opt_ptr(sync, tpl_a2b_sync_bytes(dt, B, L, &cfg), "sync", angles);
        hipStream_t cs;
        TORCH_CHECK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) == hipSuccess,
                    "StepGraph: hipStreamCreateWithFlags failed");
        std::string err;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            err = "hipStreamBeginCapture failed";
        } else {
            int rc = tpl_a2b_forward(dt, angles.data_ptr(), output_coords.data_ptr(),
                                     lengths.data_ptr<int32_t>(), st, sy, nullptr, B, L, &cfg, cs);
            if (rc == TPL_A2B_OK)
                rc = tpl_a2b_backward(dt, grad_input.data_ptr(), grad_output.data_ptr(), angles.data_ptr(),
                                      lengths.data_ptr<int32_t>(), st, sy, nullptr, B, L, &cfg, cs);
            hipGraph_t g = nullptr;
            const hipError_t ec = hipStreamEndCapture(cs, &g);
            if (rc != TPL_A2B_OK || ec != hipSuccess || !g)
                err = rc != TPL_A2B_OK ? tpl_a2b_strerror(rc) : "hipStreamEndCapture failed";
            else if (hipGraphInstantiate(&exec_, g, nullptr, nullptr, 0) != hipSuccess)
                err = "hipGraphInstantiate failed";
            if (g) (void)hipGraphDestroy(g);
        }
        (void)hipStreamDestroy(cs);
        TORCH_CHECK(err.empty(), "StepGraph: ", err);
    }
    StepGraph(const StepGraph&) = delete;
    StepGraph& operator=(const StepGraph&) = delete;
    ~StepGraph() {
        if (exec_) (void)hipGraphExecDestroy(exec_);
    }
    void replay() {
        c10::DeviceGuard guard(held_[0].device());
        hipStream_t s = c10::hip::getCurrentHIPStream(device_).stream();
        TORCH_CHECK(hipGraphLaunch(exec_, s) == hipSuccess, "StepGraph: hipGraphLaunch failed");
    }

  private:
    std::vector<torch::Tensor> held_;
    hipGraphExec_t exec_ = nullptr;
    int device_ = 0;
};

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    pybind11::class_<StepGraph>(m, "StepGraph")
        .def(pybind11::init<torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor, torch::Tensor,
                            torch::Tensor, torch::Tensor, std::vector<int64_t>>(),
             "capture one fwd+bwd step into a HIP graph")
        .def("replay", &StepGraph::replay, "launch the step on the current stream");
    m.doc() = "Angles2Backbone for MI355X: tensor glue over the tpl_a2b C ABI";
    m.def("Angles2BackboneGPU_forward", &forward, "Angles2Backbone gpu forward (chunk scan)");
    m.def("Angles2BackboneGPU_backward", &backward, "Angles2Backbone gpu backward (O(L) adjoint)");
    m.def("Angles2BackboneCPU_forward", &cpu_forward, "Angles2Backbone cpu forward (float64)");
    m.def("Angles2BackboneCPU_backward", &cpu_backward, "Angles2Backbone cpu backward (float64, O(L) adjoint)");
    m.def("Angles2BackboneGPU_forward_plain", &forward_plain, "plain serial forward");
    m.def("Angles2BackboneGPU_backward_plain", &backward_plain, "plain O(L^2) backward");
    m.def("workspace_bytes", &workspace_bytes, "(state_bytes, sync_bytes) for a config");
    m.def("sync_clear", &sync_clear, "zero a look-back buffer on the current stream");
    m.def("plain_frames_bytes", &plain_frames_bytes, "frame buffer bytes of the plain form");
    m.def("default_config", &default_config, "untuned config for (batch, max_length)");
    m.def("tune", &tune, "time candidate configs; returns (best, ms per candidate)");
    m.def("bench_steps", &bench_steps, "time back-to-back fwd+bwd steps (event ms, wall ms)");
    m.def("bench_graph", &bench_graph, "time fwd+bwd steps replayed from HIP graphs (event ms, wall ms)");
}
