// recon_tomo: slice-by-slice tomography (vertical rotation axis) from a TOML
// configuration, the reference's src/mbir2d.cpp:103-210 over this library
// (prepare_sinogram and gpu::MBIR2D of laminocam_tomocam.hpp).
//   recon_tomo <input.toml>
// Every dataset must have gamma = beta = 0; per-projection shifts are read
// but ignored, with a warning, as there.  A failed GPU reconstruction is
// reported and exits with status 1 (no CPU fallback).
#define LAMINOCAM_STANDALONE_ARRAY
#include "laminocam_tomocam.hpp"
#include "laminocam_io.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <iostream>

using namespace tomocam;

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "Usage: " << argv[0] << " <input.toml>\n";
        std::cerr << "Writing an example input file to config_template.toml.\n";
        dump_config(TOMO_CONFIG_TEMPLATE, "config_template.toml");
        return 1;
    }
    Array<float> recon;
    OutputParams output;
    try {
        const toml::Table config = read_toml_file(argv[1]);
        auto datasets = parse_input_datasets<float>(config, std::filesystem::path(argv[1]).parent_path());
        ReconParams params = parse_recon_params(config);
        output = parse_output_params(config);
        if (datasets.empty()) throw std::runtime_error("recon_tomo: no [[input]] datasets found");
        const auto &rd = params.recon_dims;
        if (rd[1] != rd[2])
            throw std::runtime_error("recon_tomo: recon_dims must be square in-plane (n2 == n3), got [" +
                                     std::to_string(rd[0]) + ", " + std::to_string(rd[1]) + ", " +
                                     std::to_string(rd[2]) + "]");
        constexpr float TILT_EPS = 1e-6f;
        for (size_t j = 0; j < datasets.size(); ++j) {
            if (std::abs(datasets[j].gamma) > TILT_EPS || std::abs(datasets[j].beta) > TILT_EPS) {
                char b[256];
                std::snprintf(b, sizeof b,
                              "recon_tomo: dataset %zu has gamma=%.4f, beta=%.4f (radians) -- this executable only "
                              "supports vertical-axis tomography (gamma=beta=0). Use recon_lamino for laminography.",
                              j, (double)datasets[j].gamma, (double)datasets[j].beta);
                throw std::runtime_error(b);
            }
            for (const auto &s : datasets[j].shifts)
                if (std::abs(s[0]) > TILT_EPS || std::abs(s[1]) > TILT_EPS) {
                    std::cerr << "Warning: dataset " << j
                              << " has non-zero per-projection shifts; recon_tomo does not yet support "
                                 "center-of-rotation shifts and will ignore them.\n";
                    break;
                }
        }
        if (params.numGpus > 1)
            std::cerr << "Warning: num_gpus = " << params.numGpus << ": this build reconstructs on one GPU\n";
        auto sino = prepare_sinogram(datasets);
        if (rd[0] != sino.data.nslices())
            std::cerr << "Warning: recon_dims[0]=" << rd[0] << " does not match the input data's row count ("
                      << sino.data.nslices() << "); reconstructing " << rd[0] << " slices from "
                      << sino.data.nslices() << "-row projections.\n";
        params.print(std::cout);
        std::cout << std::endl;

        const auto t0 = std::chrono::steady_clock::now();
        try {
            std::cout << "Running reconstruction on GPU...\n";
            recon = gpu::MBIR2D<float>(sino.data, sino.theta, params);
        } catch (const std::exception &e) {
            std::cerr << "GPU reconstruction failed: " << e.what() << "\n";
            return 1;
        }
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        char b[64];
        std::snprintf(b, sizeof b, "%.2f", s);
        std::cout << "Reconstruction completed in " << b << " seconds.\n";

        const auto base_dir = std::filesystem::path(output.filepath).parent_path();
        if (!base_dir.empty() && !std::filesystem::exists(base_dir)) std::filesystem::create_directories(base_dir);
        if (output.has_format("tiff")) {
            tiff::write(output.filepath, recon);
            std::cout << "Saved reconstruction to " << output.filepath << "\n";
        }
    } catch (const std::exception &e) {
        std::cerr << "recon_tomo: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
