// pdb_ext.cpp -- the PDB2Coords part of the pybind11 module `_FullAtomModel`:
// tensor glue over the C ABI in include/tpl_pdb.h, with the reference's
// entry-point name and argument order (Layers/FullAtomModel/main.cpp:21,
// Layers/FullAtomModel/PDB2Coords/pdb2coords_interface.cpp:65-128): CPU
// tensors, outputs resized to the largest file's atom count.
#include <torch/extension.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "tpl_pdb.h"

namespace {

// f(i) for i in [0, n) on up to torch's intra-op thread count of host threads
// (files are independent; each call parses one file).
template <typename F>
void for_files(int64_t n, F f) {
    const int64_t workers = std::max<int64_t>(1, std::min<int64_t>(n, at::get_num_threads()));
    std::atomic<int64_t> next{0};
    auto run = [&] {
        for (int64_t i = next++; i < n; i = next++) f(i);
    };
    std::vector<std::thread> pool;
    for (int64_t w = 1; w < workers; ++w) pool.emplace_back(run);
    run();
    for (auto& t : pool) t.join();
}

std::string row_string(const torch::Tensor& names, int64_t i) {
    const auto* p = names.data_ptr<uint8_t>() + i * names.size(1);
    int64_t n = 0;
    while (n < names.size(1) && p[n]) ++n;
    return std::string(reinterpret_cast<const char*>(p), n);
}

void pdb2coords_unordered(torch::Tensor filenames, torch::Tensor coords, torch::Tensor chain_names,
                          torch::Tensor res_names, torch::Tensor res_nums, torch::Tensor atom_names,
                          torch::Tensor num_atoms) {
    TORCH_CHECK(!filenames.is_cuda() && filenames.scalar_type() == torch::kByte && filenames.dim() == 2 &&
                    filenames.is_contiguous(),
                "filenames must be a contiguous CPU uint8 [batch, length] tensor");
    for (const auto* t : {&chain_names, &res_names, &atom_names})
        TORCH_CHECK(!t->is_cuda() && t->scalar_type() == torch::kByte, "names must be CPU uint8 tensors");
    TORCH_CHECK(!res_nums.is_cuda() && res_nums.scalar_type() == torch::kInt, "res_nums must be CPU int32");
    TORCH_CHECK(!num_atoms.is_cuda() && num_atoms.scalar_type() == torch::kInt && num_atoms.is_contiguous(),
                "num_atoms must be a contiguous CPU int32 tensor");
    TORCH_CHECK(!coords.is_cuda() && coords.scalar_type() == torch::kDouble && coords.dim() == 2,
                "coords must be a CPU float64 tensor of 2 dimensions");
    const int64_t B = filenames.size(0);
    TORCH_CHECK(num_atoms.numel() == B, "num_atoms must be [batch]");
    std::vector<std::string> names(B);
    for (int64_t i = 0; i < B; ++i) names[i] = row_string(filenames, i);
    std::vector<int> counts(B);
    for_files(B, [&](int64_t i) { counts[i] = tpl_pdb_count(names[i].c_str()); });
    int64_t max_atoms = 0;
    for (int64_t i = 0; i < B; ++i) {
        TORCH_CHECK(counts[i] >= 0, "PDB2CoordsUnordered: ", names[i], ": ", tpl_pdb_strerror(counts[i]));
        num_atoms.data_ptr<int32_t>()[i] = counts[i];
        max_atoms = std::max<int64_t>(max_atoms, counts[i]);
    }
    coords.resize_({B, 3 * max_atoms}).zero_();
    chain_names.resize_({B, max_atoms, 4}).zero_();
    res_names.resize_({B, max_atoms, 4}).zero_();
    res_nums.resize_({B, max_atoms}).zero_();
    atom_names.resize_({B, max_atoms, 4}).zero_();
    double* cp = coords.data_ptr<double>();
    uint8_t *chp = chain_names.data_ptr<uint8_t>(), *rnp = res_names.data_ptr<uint8_t>(),
            *anp = atom_names.data_ptr<uint8_t>();
    int32_t *nup = res_nums.data_ptr<int32_t>(), *nap = num_atoms.data_ptr<int32_t>();
    for_files(B, [&](int64_t i) {
        nap[i] = tpl_pdb_read(names[i].c_str(), cp + i * 3 * max_atoms, chp + i * 4 * max_atoms,
                              rnp + i * 4 * max_atoms, nup + i * max_atoms, anp + i * 4 * max_atoms,
                              (int)max_atoms);
    });
    for (int64_t i = 0; i < B; ++i) {
        const int n = num_atoms.data_ptr<int32_t>()[i];
        TORCH_CHECK(n >= 0, "PDB2CoordsUnordered: ", names[i], ": ", tpl_pdb_strerror(n));
    }
}

}  // namespace

void register_pdb(pybind11::module& m) {
    namespace py = pybind11;
    m.def("PDB2CoordsUnordered", &pdb2coords_unordered, "heavy-atom ATOM records of PDB files (CPU)",
          py::arg("filenames"), py::arg("coords"), py::arg("chain_names"), py::arg("res_names"),
          py::arg("res_nums"), py::arg("atom_names"), py::arg("num_atoms"));
}
