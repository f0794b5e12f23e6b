// forward_scalar: synthetic laminography projections of a scalar volume, the
// reference's src/radon.cpp:42-180 over this library (tomocam::forward).
//   forward_scalar <input.toml>
//
//   [input]
//   filename = "volume.tiff"               # {slices, rows, cols} float stack
//   [output]
//   basedir = "out"                        # created if missing
//   pad_factor = 1.5                       # volume padding before projecting
//   proj_angles = { min = -90, max = 90, n = 180 }   # degrees, [min, max) in n steps
//   [[projections]]                        # one output stack per entry
//   gamma = 30.0                           # tilt, degrees
//   filename = "gamma30.tiff"
//
// The angles (degrees) go to basedir/angles.txt.  Each stack is the forward
// projection of the volume, zero-padded symmetrically to odd sizes
// size_t(n * pad_factor) when pad_factor > 1, onto a detector of the padded
// volume's rows x cols, cropped symmetrically back to the volume's rows x
// cols, one page per angle.
#define LAMINOCAM_STANDALONE_ARRAY
#include "laminocam_tomocam.hpp"
#include "laminocam_io.hpp"

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>

using namespace tomocam;

int main(int argc, char **argv) {
    if (argc < 2) {
        std::cerr << "Usage: " << argv[0] << " <input.toml>\n";
        return 1;
    }
    try {
        const toml::Table cfg = toml::parse_file(argv[1]);
        if (!cfg.contains("input") || !cfg.contains("output")) {
            std::cerr << "Error: input and output sections are required in the TOML file.\n";
            return 1;
        }
        const toml::Value &input = cfg["input"];
        const std::string input_file = input["filename"].value_or(std::string(""));
        if (!std::filesystem::exists(input_file)) {
            std::cerr << "Error: input file " << input_file << " does not exist.\n";
            return 1;
        }
        std::cout << "Reading volume from: " << input_file << "\n";
        Array<float> volume = tiff::read(input_file);
        std::cout << "Volume loaded: " << volume.nslices() << "x" << volume.nrows() << "x" << volume.ncols()
                  << " (slices x rows x cols)\n";

        const toml::Value &output = cfg["output"];
        const std::string output_path = output["basedir"].value_or(std::string("."));
        if (!std::filesystem::exists(output_path)) std::filesystem::create_directories(output_path);
        const float pad_factor = output["pad_factor"].value_or<float>(1.0f);

        const toml::Value &ang = output["proj_angles"];
        const float ang_min = ang["min"].value_or<float>(-90.0f);
        const float ang_max = ang["max"].value_or<float>(90.0f);
        const int ang_n = ang["n"].value_or<int>(180);
        if (ang_n <= 0) throw std::runtime_error("[output] proj_angles.n must be positive");
        const float ang_step = (ang_max - ang_min) / ang_n;
        std::vector<float> angles(ang_n);
        for (int i = 0; i < ang_n; i++) angles[i] = ang_min + i * ang_step;
        std::cout << "Generating " << ang_n << " angles from " << ang_min << " to " << ang_max << " degrees\n";
        {
            std::ofstream out(std::filesystem::path(output_path) / "angles.txt");
            for (float a : angles) out << a << "\n";
        }
        std::cout << "Angles written to: " << (std::filesystem::path(output_path) / "angles.txt").string() << "\n";

        struct Projection {
            float gamma;
            std::string filename;
        };
        std::vector<Projection> projections;
        if (const toml::Array *arr = cfg["projections"].as_array()) {
            for (const toml::Value &e : *arr) {
                const toml::Table *t = e.as_table();
                if (t && t->contains("gamma") && t->contains("filename"))
                    projections.push_back({(*t)["gamma"].value_or<float>(0.0f),
                                           (*t)["filename"].value_or(std::string("proj.tiff"))});
            }
        }
        if (projections.empty()) projections.push_back({0.0f, "proj.tiff"});

        const dims_t proj_dims{(size_t)ang_n, volume.nrows(), volume.ncols()};
        auto pad = [pad_factor](size_t dim) {
            size_t n = static_cast<size_t>(dim * pad_factor);
            if (n % 2 == 0) n -= 1;
            return n;
        };
        const dims_t padded{pad(volume.nslices()), pad(volume.nrows()), pad(volume.ncols())};
        if (pad_factor > 1.0f) {
            std::cout << "Padding volume to: " << padded.n1 << "x" << padded.n2 << "x" << padded.n3 << "\n";
            volume = pad3d(volume, padded, PadType::SYMMETRIC);
            std::cout << "Padding complete\n";
        }
        std::cout << "Computing " << projections.size() << " projection(s)\n";

        int idx = 0;
        for (const auto &[gamma, filename] : projections) {
            ++idx;
            char b[64];
            std::snprintf(b, sizeof b, "%.2f", (double)gamma);
            std::cout << "[" << idx << "/" << projections.size() << "] gamma=" << b << " deg -> " << filename << "\n";
            std::vector<float> gamma_vec(angles.size(), (float)(gamma * M_PI / 180.0f));
            std::vector<float> beta_vec(angles.size(), 0.0f);
            std::vector<float> theta_vec;
            for (float a : angles) theta_vec.push_back((float)(a * M_PI / 180.0f));
            cpu::PolarGrid<float> pg(theta_vec, gamma_vec, beta_vec, volume.nrows(), volume.ncols());
            Array<float> proj = forward(volume, pg);
            proj = crop2d(proj, proj_dims, PadType::SYMMETRIC);
            const std::string out_file = (std::filesystem::path(output_path) / filename).string();
            tiff::write(out_file, proj);
            std::cout << "  Written: " << out_file << "\n";
        }
        std::cout << "Done.\n";
    } catch (const std::exception &e) {
        std::cerr << "forward_scalar: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
