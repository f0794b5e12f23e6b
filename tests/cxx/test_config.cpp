// Parses a reconstruction configuration with the C++ layer
// (include/laminocam_io.hpp) and prints what it read as JSON: per dataset the
// projection shape, a checksum of the projections, angles, gamma, beta and
// shifts (floats printed with 9 significant digits, exact for float32; the
// checksum with 17), then
// the reconstruction and output parameters.  A rejected configuration prints
// "ERROR: <message>" and exits with status 2.  tests/test_config.py compares
// the output with laminocam_amd.config on the same files.
#define LAMINOCAM_STANDALONE_ARRAY
#include "laminocam_tomocam.hpp"
#include "laminocam_io.hpp"

#include <cstdio>
#include <sstream>
#include <string>

using namespace tomocam;

static std::string num(double v, int digits = 9) {
    char b[40];
    std::snprintf(b, sizeof b, "%.*g", digits, v);
    return b;
}

int main(int argc, char **argv) {
    if (argc < 2) return 1;
    try {
        const toml::Table cfg = read_toml_file(argv[1]);
        const auto ds = parse_input_datasets<float>(cfg, std::filesystem::path(argv[1]).parent_path());
        const ReconParams p = parse_recon_params(cfg);
        const OutputParams o = parse_output_params(cfg);
        std::ostringstream js;
        js << "{\"datasets\": [";
        for (size_t d = 0; d < ds.size(); ++d) {
            const auto &x = ds[d];
            double sum = 0.0;
            for (float v : x.projs) sum += (double)v;
            js << (d ? ", " : "") << "{\"shape\": [" << x.projs.nslices() << ", " << x.projs.nrows() << ", "
               << x.projs.ncols() << "], \"sum\": " << num(sum, 17) << ", \"angles\": [";
            for (size_t i = 0; i < x.angles.size(); ++i) js << (i ? ", " : "") << num(x.angles[i]);
            js << "], \"gamma\": " << num(x.gamma) << ", \"beta\": " << num(x.beta) << ", \"shifts\": [";
            for (size_t i = 0; i < x.shifts.size(); ++i)
                js << (i ? ", " : "") << "[" << num(x.shifts[i][0]) << ", " << num(x.shifts[i][1]) << "]";
            js << "]}";
        }
        js << "], \"params\": {\"regularizer\": " << (p.regularizer == Regularizer::SPLIT_BREGMAN ? 0 : 1)
           << ", \"recon_dims\": [" << p.recon_dims[0] << ", " << p.recon_dims[1] << ", " << p.recon_dims[2]
           << "], \"maxIters\": " << p.maxIters << ", \"innerIters\": " << p.innerIters
           << ", \"lambda\": " << num(p.lambda) << ", \"mu\": " << num(p.mu) << ", \"tol\": " << num(p.tol)
           << ", \"xtol\": " << num(p.xtol) << ", \"numGpus\": " << p.numGpus
           << ", \"chunkSlices\": " << p.chunkSlices << ", \"streamed\": " << (p.streamed ? "true" : "false")
           << "}, \"output\": {\"filepath\": \"" << o.filepath << "\", \"formats\": [";
        for (size_t i = 0; i < o.formats.size(); ++i) js << (i ? ", " : "") << "\"" << o.formats[i] << "\"";
        js << "]}}";
        std::printf("%s\n", js.str().c_str());
    } catch (const std::exception &e) {
        std::printf("ERROR: %s\n", e.what());
        return 2;
    }
    return 0;
}
