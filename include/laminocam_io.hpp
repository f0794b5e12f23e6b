/*
 * laminocam_io.hpp - the reference's file-format, padding and configuration
 * layer of its reconstruction programs, header only, over liblaminocam.so's
 * host codecs (lc_tiff_* / lc_h5_* of laminocam.h) and laminocam_toml.hpp.
 * Include laminocam_tomocam.hpp first.
 *
 *   include/tiff.h:31-117        tiff::read, tiff::write
 *   include/hdfread.h:34-96      h5::read_images, h5::read_angles
 *   include/mask.h:43-112        mask_infs_nans, mask_support
 *   include/padding.h:33-196     PadType, pad2d, crop2d, pad3d, crop3d
 *   include/config.h:45-489      read_toml_file, resolve_path, warn_unknown_keys,
 *                                angles_to_radians, read_angles_file,
 *                                read_shifts_file, parse_input_datasets,
 *                                parse_recon_params, warn_laminography_dims,
 *                                parse_output_params, dump_config
 *
 * The configuration keys, defaults, validation and error messages are the
 * reference's (its [[input]], [output] and [recon_params] tables); toml++ is
 * replaced by laminocam::toml (the same node semantics).  The example
 * configurations written by dump_config are this library's own text.
 */
#ifndef LAMINOCAM_IO_HPP
#define LAMINOCAM_IO_HPP

#include <algorithm>
#include <array>
#include <cctype>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <initializer_list>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "laminocam.h"
#include "laminocam_toml.hpp"

#ifndef LAMINOCAM_TOMOCAM_HPP
#error "include laminocam_tomocam.hpp before laminocam_io.hpp"
#endif

namespace tomocam {

namespace toml = laminocam::toml;

// ---------------------------------------------------------------------------
// padding (include/padding.h)
// ---------------------------------------------------------------------------
#ifdef LAMINOCAM_STANDALONE_ARRAY
enum class PadType { LEFT, RIGHT, SYMMETRIC };
#endif

// slices padded to size_t(factor * n) in each in-plane dimension, made odd
template <typename T> Array<T> pad2d(const Array<T> &arr, float factor, PadType pad_type) {
    if (factor - 1 < 1.e-06) return arr;
    size_t n2 = static_cast<size_t>(factor * arr.nrows());
    if (n2 % 2 == 0) n2 -= 1;
    size_t n3 = static_cast<size_t>(factor * arr.ncols());
    if (n3 % 2 == 0) n3 -= 1;
    Array<T> out(dims_t{arr.nslices(), n2, n3});
    std::fill(out.begin(), out.end(), T(0));
    size_t d1 = (n2 - arr.nrows()) / 2, d2 = (n3 - arr.ncols()) / 2;
    if (pad_type == PadType::RIGHT) d1 = d2 = 0;
    else if (pad_type == PadType::LEFT) d1 = n2 - arr.nrows(), d2 = n3 - arr.ncols();
    for (size_t i = 0; i < arr.nslices(); ++i)
        for (size_t j = 0; j < arr.nrows(); ++j)
            std::copy(arr.begin() + (i * arr.nrows() + j) * arr.ncols(),
                      arr.begin() + (i * arr.nrows() + j + 1) * arr.ncols(),
                      out.begin() + (i * n2 + j + d1) * n3 + d2);
    return out;
}

// each slice cropped to new_dims.n2 x new_dims.n3 (LEFT keeps the first rows
// and columns, RIGHT the last, SYMMETRIC the centre)
template <typename T> Array<T> crop2d(const Array<T> &arr, dims_t new_dims, PadType pad_type) {
    if (new_dims.n1 != arr.nslices()) throw std::runtime_error("crop2d: new_dims.n1 must equal arr.nslices()");
    const size_t c2 = arr.nrows() - new_dims.n2, c3 = arr.ncols() - new_dims.n3;
    if (c2 == 0 && c3 == 0) return arr;
    Array<T> out(new_dims);
    size_t d1 = c2 / 2, d2 = c3 / 2;
    if (pad_type == PadType::LEFT) d1 = d2 = 0;
    else if (pad_type == PadType::RIGHT) d1 = c2, d2 = c3;
    for (size_t i = 0; i < new_dims.n1; ++i)
        for (size_t j = 0; j < new_dims.n2; ++j)
            std::copy(arr.begin() + (i * arr.nrows() + j + d1) * arr.ncols() + d2,
                      arr.begin() + (i * arr.nrows() + j + d1) * arr.ncols() + d2 + new_dims.n3,
                      out.begin() + (i * new_dims.n2 + j) * new_dims.n3);
    return out;
}

// the volume zero-padded to dims (RIGHT: at the start, LEFT: at the end of
// each dimension, SYMMETRIC: centred, the extra element at the end)
template <typename T> Array<T> pad3d(const Array<T> &arr, dims_t dims, PadType pad_type) {
    Array<T> out(dims);
    std::fill(out.begin(), out.end(), T(0));
    dims_t d{(dims.n1 - arr.nslices()) / 2, (dims.n2 - arr.nrows()) / 2, (dims.n3 - arr.ncols()) / 2};
    if (pad_type == PadType::RIGHT) d = dims_t{0, 0, 0};
    else if (pad_type == PadType::LEFT)
        d = dims_t{dims.n1 - arr.nslices(), dims.n2 - arr.nrows(), dims.n3 - arr.ncols()};
    for (size_t i = 0; i < arr.nslices(); ++i)
        for (size_t j = 0; j < arr.nrows(); ++j)
            std::copy(arr.begin() + (i * arr.nrows() + j) * arr.ncols(),
                      arr.begin() + (i * arr.nrows() + j + 1) * arr.ncols(),
                      out.begin() + ((i + d.n1) * dims.n2 + j + d.n2) * dims.n3 + d.n3);
    return out;
}

template <typename T> Array<T> crop3d(const Array<T> &arr, dims_t new_dims, PadType pad_type) {
    const dims_t c{arr.nslices() - new_dims.n1, arr.nrows() - new_dims.n2, arr.ncols() - new_dims.n3};
    if (c.n1 == 0 && c.n2 == 0 && c.n3 == 0) return arr;
    dims_t d{c.n1 / 2, c.n2 / 2, c.n3 / 2};
    if (pad_type == PadType::LEFT) d = dims_t{0, 0, 0};
    else if (pad_type == PadType::RIGHT) d = c;
    Array<T> out(new_dims);
    for (size_t i = 0; i < new_dims.n1; ++i)
        for (size_t j = 0; j < new_dims.n2; ++j)
            std::copy(arr.begin() + ((i + d.n1) * arr.nrows() + j + d.n2) * arr.ncols() + d.n3,
                      arr.begin() + ((i + d.n1) * arr.nrows() + j + d.n2) * arr.ncols() + d.n3 + new_dims.n3,
                      out.begin() + (i * new_dims.n2 + j) * new_dims.n3);
    return out;
}

// ---------------------------------------------------------------------------
// TIFF and HDF5 (include/tiff.h, include/hdfread.h)
// ---------------------------------------------------------------------------
namespace tiff {
// every page; even heights and widths lose their last row / column
inline Array<float> read(std::string filename) {
    int64_t n = 0, h = 0, w = 0;
    laminocam::check(lc_tiff_info(filename.c_str(), &n, &h, &w));
    if (h % 2 == 0) --h;
    if (w % 2 == 0) --w;
    Array<float> data(dims_t{(size_t)n, (size_t)h, (size_t)w});
    laminocam::check(lc_tiff_read(filename.c_str(), data.data(), n, h, w));
    return data;
}

inline void write(std::string filename, const Array<float> &data) {
    laminocam::check(lc_tiff_write(filename.c_str(), data.data(), (int64_t)data.nslices(), (int64_t)data.nrows(),
                                   (int64_t)data.ncols()));
}
} // namespace tiff

namespace h5 {
// a 3-D dataset {nangles, nrows, ncols} as float
inline Array<float> read_images(const std::string &filename, const std::string &dataset = "/images") {
    int32_t rank = 0;
    int64_t dims[32] = {};
    laminocam::check(lc_h5_shape(filename.c_str(), dataset.c_str(), &rank, dims));
    if (rank != 3) throw std::runtime_error("expected 3-D dataset: " + dataset);
    Array<float> out(dims_t{(size_t)dims[0], (size_t)dims[1], (size_t)dims[2]});
    laminocam::check(lc_h5_read(filename.c_str(), dataset.c_str(), out.data(), (int64_t)out.size(), 0));
    return out;
}

// a 1-D dataset as floats
inline std::vector<float> read_angles(const std::string &filename, const std::string &dataset = "/coords/alpha") {
    int32_t rank = 0;
    int64_t dims[32] = {};
    laminocam::check(lc_h5_shape(filename.c_str(), dataset.c_str(), &rank, dims));
    if (rank != 1) throw std::runtime_error("expected 1-D dataset: " + dataset);
    std::vector<float> out((size_t)dims[0]);
    laminocam::check(lc_h5_read(filename.c_str(), dataset.c_str(), out.data(), dims[0], 0));
    return out;
}
} // namespace h5

// ---------------------------------------------------------------------------
// masks (include/mask.h)
// ---------------------------------------------------------------------------
template <typename T> Array<T> mask_infs_nans(const Array<T> &projs) {
    Array<T> masked = projs;
    for (T &v : masked)
        if (!std::isfinite(v)) v = 0;
    return masked;
}

// detector pixels outside the rectangular support of a dims-sized object, at
// tilt gamma, set to 0
template <typename T>
Array<T> mask_support(const Array<T> &projs, const dims_t dims, const std::vector<T> &theta, const T gamma) {
    const T cos_g = std::cos(gamma), sin_g = std::sin(gamma);
    const T xcen = projs.ncols() / 2, ycen = projs.nrows() / 2;
    const T ylim = dims.n2 / 2, xlim = dims.n3 / 2;
    Array<T> mask = projs;
    for (size_t p = 0; p < theta.size(); ++p) {
        const T cos_t = std::cos(theta[p]);
        if (std::abs(cos_t) < 1e-06)
            throw std::runtime_error("Projection angle too close to 90 degrees, support mask is undefined.");
        for (size_t y = 0; y < projs.nrows(); ++y)
            for (size_t x = 0; x < projs.ncols(); ++x) {
                const T xr = x - xcen, yr = y - ycen;
                const T xc = (xr * cos_g * cos_t + yr * sin_g) / cos_t;
                const T yc = (-xr * sin_g * cos_t + yr * cos_g) / cos_t;
                if (std::abs(xc) > xlim || std::abs(yc) > ylim) mask[(p * projs.nrows() + y) * projs.ncols() + x] = 0;
            }
    }
    return mask;
}

// ---------------------------------------------------------------------------
// configuration (include/config.h)
// ---------------------------------------------------------------------------
inline toml::Table read_toml_file(const std::string &filepath) {
    if (!std::filesystem::exists(filepath)) throw std::runtime_error("TOML file does not exist: " + filepath);
    try {
        return toml::parse_file(filepath);
    } catch (const toml::parse_error &err) {
        throw std::runtime_error("Failed to parse TOML file '" + filepath + "': " + err.what());
    }
}

// a relative path is taken relative to the config file's directory when the
// file exists there, else as given
inline std::string resolve_path(const std::string &p, const std::filesystem::path &base_dir) {
    std::filesystem::path fp(p);
    if (fp.is_relative() && !base_dir.empty()) {
        const auto candidate = base_dir / fp;
        if (std::filesystem::exists(candidate)) return candidate.string();
    }
    return p;
}

inline void warn_unknown_keys(const toml::Table &tbl, std::initializer_list<const char *> allowed,
                              const std::string &section) {
    for (const std::string &key : tbl.keys()) {
        bool known = false;
        for (const char *a : allowed) known = known || key == a;
        if (!known) std::cerr << "\033[33mWarning\033[0m: unknown key '" << key << "' in " << section << " (ignored)\n";
    }
}

template <typename T> void angles_to_radians(std::vector<T> &angles, const std::string &units) {
    if (units == "deg") {
        for (auto &a : angles) a = a * T(M_PI) / T(180);
    } else if (units != "rad") {
        throw std::runtime_error("[[input]] 'angle_units' must be \"deg\" or \"rad\", got '" + units + "'");
    }
}

// one angle per whitespace-separated value, in `units`, returned in radians
template <typename T> std::vector<T> read_angles_file(const std::string &filepath, const std::string &units = "deg") {
    std::ifstream fp(filepath);
    if (!fp.is_open()) throw std::runtime_error("Could not open angles file: " + filepath);
    std::vector<T> angles;
    T a;
    while (fp >> a) angles.push_back(a);
    if (angles.empty()) throw std::runtime_error("No angles found in file: " + filepath);
    angles_to_radians(angles, units);
    return angles;
}

// pairs "dx dy" (pixels), one per projection
template <typename T> std::vector<std::array<T, 2>> read_shifts_file(const std::string &filepath) {
    std::ifstream fp(filepath);
    if (!fp.is_open()) throw std::runtime_error("Could not open shifts file: " + filepath);
    std::vector<std::array<T, 2>> shifts;
    T dx, dy;
    while (fp >> dx >> dy) shifts.push_back({dx, dy});
    if (shifts.empty()) throw std::runtime_error("No shifts found in file: " + filepath);
    return shifts;
}

namespace detail_io {
inline bool exists(const std::string &p) { return std::filesystem::exists(p); }
template <typename T> T number(const toml::Value &v, T d) {
    auto x = v.value<T>();
    return x ? *x : d;
}
} // namespace detail_io

// include/config.h:160-331: the [[input]] datasets.  Projections are TIFF
// stacks (with a text file of angles) or HDF5 datasets (projs_dataset,
// default /images; angles dataset, default /coords/alpha); non-finite values
// become 0; gamma (degrees, sign flipped) and beta (degrees) default to 0;
// shifts from a two-column file or a broadcast 'cor-offset' pair; an
// 'alignment' TOML replaces gamma, beta, angles and shifts.
template <typename T>
std::vector<Dataset_t<T>> parse_input_datasets(const toml::Table &config, const std::filesystem::path &base_dir = {}) {
    const toml::Array *input_array = config["input"].as_array();
    if (!input_array) throw std::runtime_error("Missing [[input]] array in TOML file");
    std::vector<Dataset_t<T>> datasets;
    for (const toml::Value &elem : *input_array) {
        const toml::Table *in = elem.as_table();
        if (!in) throw std::runtime_error("Invalid [[input]] entry");
        if (!in->contains("filename")) throw std::runtime_error("[[input]] entry must have a 'filename' field");
        warn_unknown_keys(*in, {"filename", "angles", "angle_units", "gamma", "beta", "shifts", "cor-offset",
                                "alignment", "projs_dataset"},
                          "[[input]]");
        auto filename = (*in)["filename"].value<std::string>();
        if (!filename) throw std::runtime_error("[[input]] 'filename' must be a string");
        *filename = resolve_path(*filename, base_dir);
        const std::string angle_units = (*in)["angle_units"].value_or(std::string("deg"));
        if (!detail_io::exists(*filename)) throw std::runtime_error("Projection file does not exist: " + *filename);
        T gamma_rad = -detail_io::number<T>((*in)["gamma"], T(0)) * T(M_PI) / T(180);
        T beta_rad = detail_io::number<T>((*in)["beta"], T(0)) * T(M_PI) / T(180);

        std::vector<std::array<T, 2>> shifts;
        if (in->contains("shifts")) {
            auto sp = (*in)["shifts"].value<std::string>();
            if (!sp) throw std::runtime_error("[[input]] 'shifts' must be a string path");
            const std::string resolved = resolve_path(*sp, base_dir);
            if (!detail_io::exists(resolved)) throw std::runtime_error("Shifts file does not exist: " + resolved);
            shifts = read_shifts_file<T>(resolved);
        } else if (in->contains("cor-offset")) {
            const toml::Array *off = (*in)["cor-offset"].as_array();
            if (!off || off->size() != 2) throw std::runtime_error("[[input]] 'cor-offset' must be in form [dx, dy]");
            std::array<T, 2> buf{};
            for (size_t i = 0; i < 2; ++i) {
                auto v = (*off)[i].value<T>();
                if (!v) throw std::runtime_error("[[input]] 'cor-offset' must be an array of numbers");
                buf[i] = *v;
            }
            shifts = {buf};   // one pair: broadcast once the projection count is known
        }

        std::string ext = std::filesystem::path(*filename).extension().string();
        std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return (char)std::tolower(c); });
        Array<float> projs;
        std::vector<T> angles;
        if (ext == ".tiff" || ext == ".tif") {
            auto ap = (*in)["angles"].value<std::string>();
            if (!ap) throw std::runtime_error("[[input]] 'angles' is required for TIFF files");
            *ap = resolve_path(*ap, base_dir);
            if (!detail_io::exists(*ap)) throw std::runtime_error("Angles file does not exist: " + *ap);
            projs = tiff::read(*filename);
            angles = read_angles_file<T>(*ap, angle_units);
            if (projs.nslices() != angles.size())
                throw std::runtime_error("'" + *filename + "' has " + std::to_string(projs.nslices()) + " pages but '" +
                                         *ap + "' has " + std::to_string(angles.size()) + " entries");
        } else if (ext == ".h5" || ext == ".hdf5") {
            const std::string angles_ds = (*in)["angles"].value_or(std::string("/coords/alpha"));
            const std::string projs_ds = (*in)["projs_dataset"].value_or(std::string("/images"));
            projs = h5::read_images(*filename, projs_ds);
            const std::vector<float> raw = h5::read_angles(*filename, angles_ds);
            angles.assign(raw.begin(), raw.end());
            angles_to_radians(angles, angle_units);
            if (projs.nslices() != angles.size())
                throw std::runtime_error("'" + *filename + "' images has " + std::to_string(projs.nslices()) +
                                         " slices but angles dataset '" + angles_ds + "' has " +
                                         std::to_string(angles.size()) + " entries");
        } else {
            throw std::runtime_error("Unsupported file extension '" + ext + "': " + *filename);
        }
        projs = mask_infs_nans(projs);
        if (shifts.size() == 1) {
            const std::array<T, 2> s = shifts[0];
            shifts.assign(angles.size(), s);
        } else if (shifts.empty()) {
            shifts.assign(angles.size(), std::array<T, 2>{T(0), T(0)});
        } else if (shifts.size() != angles.size()) {
            throw std::runtime_error("shifts file has " + std::to_string(shifts.size()) + " entries but " +
                                     std::to_string(angles.size()) + " projections were loaded");
        }

        if (in->contains("alignment")) {
            auto al = (*in)["alignment"].value<std::string>();
            if (!al) throw std::runtime_error("[[input]] 'alignment' must be a string path");
            *al = resolve_path(*al, base_dir);
            if (!detail_io::exists(*al)) throw std::runtime_error("Alignment file does not exist: " + *al);
            const toml::Table at = read_toml_file(*al);
            if (auto gv = at["gamma_deg"].value<T>()) gamma_rad = -(*gv) * T(M_PI) / T(180);
            if (auto bv = at["beta_deg"].value<T>()) beta_rad = (*bv) * T(M_PI) / T(180);
            const toml::Array *aa = at["corrected_angles_deg"].as_array();
            if (!aa) throw std::runtime_error("Alignment TOML missing 'corrected_angles_deg'");
            angles.clear();
            for (const toml::Value &e : *aa) {
                auto v = e.value<T>();
                if (!v) throw std::runtime_error("Alignment TOML 'corrected_angles_deg' must hold numbers");
                angles.push_back(*v * T(M_PI) / T(180));
            }
            if (const toml::Array *spx = at["shifts_px"].as_array()) {
                shifts.clear();
                for (const toml::Value &e : *spx) {
                    const toml::Array *pair = e.as_array();
                    if (!pair || pair->size() < 2)
                        throw std::runtime_error("Alignment TOML 'shifts_px' entries must be [dx, dy] pairs");
                    auto dx = (*pair)[0].value<T>(), dy = (*pair)[1].value<T>();
                    if (!dx || !dy) throw std::runtime_error("Alignment TOML 'shifts_px' entries must be [dx, dy] pairs");
                    shifts.push_back({*dx, *dy});
                }
            }
            if (shifts.size() != angles.size())
                throw std::runtime_error("Alignment TOML: shifts_px has " + std::to_string(shifts.size()) +
                                         " entries but corrected_angles_deg has " + std::to_string(angles.size()));
        }

        Dataset_t<T> ds;
        if constexpr (std::is_same_v<T, float>) {
            ds.projs = std::move(projs);
        } else {
            ds.projs = Array<T>(projs.dims());
            std::copy(projs.begin(), projs.end(), ds.projs.begin());
        }
        ds.angles = std::move(angles);
        ds.gamma = gamma_rad;
        ds.beta = beta_rad;
        ds.shifts = std::move(shifts);
        datasets.push_back(std::move(ds));
    }
    return datasets;
}

// include/config.h:333-400: [recon_params]; each recon_dims entry must be a
// positive integer and an even one is reduced by 1
inline ReconParams parse_recon_params(const toml::Table &config) {
    ReconParams p;
    const toml::Table *recon = config["recon_params"].as_table();
    if (!recon) throw std::runtime_error("Missing [recon_params] section in config file");
    warn_unknown_keys(*recon, {"max_iters", "tol", "xtol", "recon_dims", "regularizer", "num_gpus", "chunk_slices",
                               "streamed"},
                      "[recon_params]");
    p.maxIters = (*recon)["max_iters"].value_or<size_t>(50);
    const toml::Array *dims = (*recon)["recon_dims"].as_array();
    if (!dims || dims->size() != 3)
        throw std::runtime_error("[recon_params] 'recon_dims' must be an array of three integers");
    for (size_t i = 0; i < 3; ++i) {
        size_t v = (*dims)[i].value_or<size_t>(0);
        if (v == 0)
            throw std::runtime_error("[recon_params] 'recon_dims[" + std::to_string(i) + "]' must be a positive integer");
        if (v % 2 == 0) v -= 1;
        p.recon_dims[i] = v;
    }
    p.tol = (*recon)["tol"].value_or<float>(1e-5f);
    p.xtol = (*recon)["xtol"].value_or<float>(1e-5f);
    p.numGpus = (*recon)["num_gpus"].value_or<size_t>(0);
    p.chunkSlices = (*recon)["chunk_slices"].value_or<size_t>(0);
    p.streamed = (*recon)["streamed"].value_or<bool>(false);
    if (recon->contains("regularizer")) {
        if (const toml::Table *reg = (*recon)["regularizer"].as_table()) {
            warn_unknown_keys(*reg, {"method", "split_bregman"}, "[recon_params.regularizer]");
            const std::string method = (*reg)["method"].value_or(std::string("split_bregman"));
            if (method != "split_bregman")
                throw std::runtime_error("[recon_params] 'regularizer' must be 'split_bregman'");
            p.regularizer = Regularizer::SPLIT_BREGMAN;
            const toml::Table *sb = (*reg)["split_bregman"].as_table();
            if (!sb)
                throw std::runtime_error("Missing [recon_params.regularizer.split_bregman] section in config file");
            warn_unknown_keys(*sb, {"lambda", "mu", "inner_iters"}, "[recon_params.regularizer.split_bregman]");
            p.lambda = (*sb)["lambda"].value_or<float>(0.1f);
            p.mu = (*sb)["mu"].value_or<float>(10.0f);
            p.innerIters = (*sb)["inner_iters"].value_or<size_t>(1);
        }
    }
    return p;
}

// include/config.h:402-417
inline void warn_laminography_dims(const ReconParams &p) {
    const float ratio = static_cast<float>(p.recon_dims[0]) / static_cast<float>(p.recon_dims[2]);
    if (ratio > 0.15f) {
        char b[64];
        std::snprintf(b, sizeof b, "%.2f", ratio);
        std::cerr << "\033[31mWarning\033[0m: recon_dims[0] / recon_dims[2] = " << b << " > 0.15.\n"
                  << "laminography thickness (recon_dims[0]) is expected to be much smaller than the in-plane "
                     "dimensions (recon_dims[1], recon_dims[2]).\n";
    }
}

// include/config.h:419-475: [output] filename (default ./recon.tiff) and
// formats ("tiff" and / or "vti", default tiff, duplicates dropped)
inline OutputParams parse_output_params(const toml::Table &config) {
    OutputParams params;
    const toml::Value &output = config["output"];
    if (!output) throw std::runtime_error("Missing [output] section in config file");
    if (const toml::Table *t = output.as_table()) warn_unknown_keys(*t, {"filename", "formats"}, "[output]");
    params.filepath = output["filename"].value_or(std::string("./recon.tiff"));
    std::vector<std::string> formats;
    if (const toml::Array *fa = output["formats"].as_array()) {
        for (const toml::Value &e : *fa) {
            auto f = e.value<std::string>();
            if (!f) throw std::runtime_error("[output] 'formats' array must contain strings");
            if (*f != "tiff" && *f != "vti")
                throw std::runtime_error("[output] invalid format '" + *f + "'. Must be 'tiff' or 'vti'");
            formats.push_back(*f);
        }
    } else {
        formats = {"tiff"};
    }
    for (const std::string &f : formats)
        if (std::find(params.formats.begin(), params.formats.end(), f) == params.formats.end())
            params.formats.push_back(f);
    return params;
}

inline void dump_config(const char *tmpl, const std::string &filepath) {
    std::ofstream out(filepath);
    if (!out.is_open()) throw std::runtime_error("Could not open file for writing: " + filepath);
    out << tmpl;
}

// example configurations for recon_lamino and recon_tomo
inline const char *const LAMINO_CONFIG_TEMPLATE = R"(# recon_lamino: laminography reconstruction of one or more tilt series.
# Run: recon_lamino config.toml

[[input]]
filename = "projections.tiff"     # a .tif/.tiff stack (one page per projection) or .h5/.hdf5
angles = "angles.txt"             # TIFF: text file of angles; HDF5: angles dataset (default /coords/alpha)
angle_units = "deg"               # "deg" (default) or "rad"
gamma = 30.0                      # tilt of the rotation axis, degrees
# beta = 0.0                      # second tilt, degrees
# shifts = "shifts.txt"           # per-projection "dx dy" in pixels
# cor-offset = [0.0, 0.0]         # one shift for every projection
# alignment = "alignment.toml"    # gamma_deg, beta_deg, corrected_angles_deg, shifts_px
# projs_dataset = "/images"       # HDF5 projections dataset

[output]
filename = "recon.tiff"
formats = ["tiff"]

[recon_params]
max_iters = 50
tol = 1e-5
xtol = 1e-5
recon_dims = [41, 511, 511]       # (slices, y, x); even sizes are reduced by one

# omit this table for an unregularised least-squares solve
[recon_params.regularizer]
method = "split_bregman"

[recon_params.regularizer.split_bregman]
lambda = 0.1
mu = 10.0
inner_iters = 1
)";

inline const char *const TOMO_CONFIG_TEMPLATE = R"(# recon_tomo: slice-by-slice tomography about a vertical axis.
# Run: recon_tomo config.toml

[[input]]
filename = "projections.h5"       # .h5/.hdf5 or a .tif/.tiff stack
# angles = "angles.txt"           # TIFF: text file of angles (required)
# projs_dataset = "/images"       # HDF5 projections dataset (default)
# angles = "/coords/alpha"        # HDF5 angles dataset (default)
angle_units = "deg"               # "deg" (default) or "rad"
# gamma and beta must be 0 here; use recon_lamino for tilted axes

[output]
filename = "recon2d.tiff"
formats = ["tiff"]

[recon_params]
max_iters = 50
tol = 1e-5
xtol = 1e-5
recon_dims = [511, 511, 511]      # (slices, y, x); slices = detector rows; y = x

[recon_params.regularizer]
method = "split_bregman"

[recon_params.regularizer.split_bregman]
lambda = 0.1
mu = 10.0
inner_iters = 1
)";

} // namespace tomocam

#endif // LAMINOCAM_IO_HPP
