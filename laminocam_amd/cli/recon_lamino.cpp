// recon_lamino: laminography reconstruction from a TOML configuration, the
// reference's src/recon.cpp:40-98 over this library (gpu::MBIR of
// laminocam_tomocam.hpp, the configuration and file layer of laminocam_io.hpp).
//   recon_lamino <input.toml>
// With no argument it prints the usage and writes an example configuration to
// config_template.toml (exit status 1).
#define LAMINOCAM_STANDALONE_ARRAY
#include "laminocam_tomocam.hpp"
#include "laminocam_io.hpp"

#include <chrono>
#include <cstdio>
#include <iostream>

using namespace tomocam;

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "Usage: " << argv[0] << " <input.toml>\n";
        std::cerr << "Writing an example input file to config_template.toml.\n";
        dump_config(LAMINO_CONFIG_TEMPLATE, "config_template.toml");
        return 1;
    }
    try {
        const toml::Table config = read_toml_file(argv[1]);
        auto datasets = parse_input_datasets<float>(config, std::filesystem::path(argv[1]).parent_path());
        ReconParams params = parse_recon_params(config);
        warn_laminography_dims(params);
        const OutputParams output = parse_output_params(config);
        if (params.numGpus > 1)
            std::cerr << "Warning: num_gpus = " << params.numGpus << ": this build reconstructs on one GPU\n";
        params.print(std::cout);
        std::cout << std::endl;

        const auto t0 = std::chrono::steady_clock::now();
        std::cout << "Running reconstruction on GPU...\n";
        Array<float> recon = gpu::MBIR<float>(datasets, params);
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
        std::cerr << "recon_lamino: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
