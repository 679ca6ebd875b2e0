// ss_harness.cpp -- command-line gate and timing harness for the operator.
//
// Mirrors the reference's unified harness (harness/primary/gpu_common_v0_6.cu:
// CLI :500-535, per-image loop :584-745, timing protocol :305-397, CSV columns
// :545-553, percentile rule :86-103, byte mask compare :105-121, exit codes
// :752-755 and :1068-1076 of snapshot_sparse_seam_cuda_v1_0.cu) on top of the
// C ABI.  Images are binary PGM (P5) or PNG (8-bit gray, gray + alpha,
// palette, RGB or RGBA; colour converted to gray with the ITU-R 601 weights
// in 15-bit fixed point, alpha dropped); OpenCV is not used.
//
//   ss_harness <image_dir> --output=DIR [--expected-max=DIR --expected-min=DIR]
//              [--warmup=5] [--reps=30] [--tile=T --threads=N | --autotune]
//              [--levels=L] [--device=D] [--skip-expected]
//
// Every image is also run through the dense Global-UF route
// (ss_global_uf_u8) and its masks compared with the operator's: the
// cross_*_mismatches columns (gpu_common_v0_6.cu:659-669).
// Exit codes: 0 all masks match (oracle and cross-route), 1 mismatches,
// 2 usage, 3 error.
//
//   ss_harness --convert <image_dir> --output=DIR
//
// Writes every input as the 8-bit gray PGM the other modes read (no GPU).
//
//   ss_harness --seam-profile <image_dir> --output=FILE [--tiles=32,64,128]
//              [--levels=256,32,8,2] [--device=D]
//
// Seam-sparsity mode, mirroring the reference's profiler
// (harness/seam_sparsity_profiler/gpu_seam_sparsity_profiler_v0_1.cu: CLI
// :341-383, per image x levels x tiles rows :391-517, exit codes :518-540):
// one CSV row per (image, levels, tile) with the same columns.  Tiles are
// 16, 32, 64 or 128 (default 32, 64, 128 as in the reference).  The reference quantises to bin
// indices ((v L) >> 8) there; the rescaled quantiser used here maps the same
// bins to distinct values, so every count is the same.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include <zlib.h>

#include "sparse_seam.hpp"

namespace fs = std::filesystem;

namespace {

struct Pgm {
    int width = 0, height = 0;
    std::vector<uint8_t> data;
};

// Binary PGM reader; the header is tokenised, so a first pixel byte that is
// itself whitespace is not swallowed.
Pgm read_pgm(const fs::path& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot read input " + path.string());
    std::vector<char> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    size_t pos = 0;
    std::vector<std::string> tok;
    while (tok.size() < 4) {
        while (pos < buf.size() && std::isspace((unsigned char)buf[pos])) ++pos;
        if (pos < buf.size() && buf[pos] == '#') {
            while (pos < buf.size() && buf[pos] != '\n') ++pos;
            continue;
        }
        const size_t s = pos;
        while (pos < buf.size() && !std::isspace((unsigned char)buf[pos])) ++pos;
        if (s == pos) throw std::runtime_error("Truncated PGM header " + path.string());
        tok.emplace_back(buf.data() + s, pos - s);
    }
    ++pos;
    if (tok[0] != "P5" || std::stoi(tok[3]) != 255) throw std::runtime_error("Unsupported PGM " + path.string());
    Pgm p;
    p.width = std::stoi(tok[1]);
    p.height = std::stoi(tok[2]);
    const size_t n = (size_t)p.width * p.height;
    if (buf.size() < pos + n) throw std::runtime_error("Truncated PGM data " + path.string());
    p.data.assign(buf.begin() + pos, buf.begin() + pos + n);
    return p;
}

uint32_t be32(const uint8_t* p) { return (uint32_t)p[0] << 24 | (uint32_t)p[1] << 16 | (uint32_t)p[2] << 8 | p[3]; }

// 8-bit gray from R, G, B: 0.299, 0.587, 0.114 in 15-bit fixed point.
uint8_t rgb_gray(uint32_t r, uint32_t g, uint32_t b) {
    return (uint8_t)((r * 9798u + g * 19235u + b * 3735u + 16384u) >> 15);
}

// PNG reader: chunks (CRC checked), zlib inflate of the IDAT stream, the five
// row filters, non-interlaced 8-bit images of colour types 0, 2, 3, 4, 6.
Pgm read_png(const fs::path& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Cannot read input " + path.string());
    const std::vector<uint8_t> buf((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    static const uint8_t sig[8] = {0x89, 'P', 'N', 'G', 0x0D, 0x0A, 0x1A, 0x0A};
    if (buf.size() < 8 || std::memcmp(buf.data(), sig, 8) != 0) throw std::runtime_error("Not a PNG " + path.string());
    int w = 0, h = 0, depth = 0, ctype = -1;
    std::vector<uint8_t> idat, plte;
    size_t pos = 8;
    bool end = false;
    while (!end) {
        if (pos + 12 > buf.size()) throw std::runtime_error("Truncated PNG " + path.string());
        const uint32_t len = be32(&buf[pos]);
        if (len > buf.size() - pos - 12) throw std::runtime_error("Truncated PNG chunk " + path.string());
        const uint8_t* type = &buf[pos + 4];
        const uint8_t* data = &buf[pos + 8];
        if ((uint32_t)crc32(crc32(0L, Z_NULL, 0), type, len + 4) != be32(data + len))
            throw std::runtime_error("PNG chunk CRC mismatch " + path.string());
        if (!std::memcmp(type, "IHDR", 4)) {
            if (len != 13) throw std::runtime_error("Bad PNG header " + path.string());
            w = (int)be32(data);
            h = (int)be32(data + 4);
            depth = data[8];
            ctype = data[9];
            if (data[10] != 0 || data[11] != 0 || data[12] != 0)
                throw std::runtime_error("Unsupported PNG (interlaced or non-standard) " + path.string());
        } else if (!std::memcmp(type, "PLTE", 4)) {
            plte.assign(data, data + len);
        } else if (!std::memcmp(type, "IDAT", 4)) {
            idat.insert(idat.end(), data, data + len);
        } else if (!std::memcmp(type, "IEND", 4)) {
            end = true;
        }
        pos += 12 + len;
    }
    const int ch = ctype == 0 ? 1 : ctype == 2 ? 3 : ctype == 3 ? 1 : ctype == 4 ? 2 : ctype == 6 ? 4 : 0;
    if (w <= 0 || h <= 0 || depth != 8 || ch == 0 || (ctype == 3 && plte.size() < 3))
        throw std::runtime_error("Unsupported PNG (8-bit gray, palette, RGB or alpha only) " + path.string());
    const size_t stride = (size_t)w * ch;
    std::vector<uint8_t> raw((stride + 1) * (size_t)h);
    uLongf rlen = (uLongf)raw.size();
    if (uncompress(raw.data(), &rlen, idat.data(), (uLong)idat.size()) != Z_OK || rlen != raw.size())
        throw std::runtime_error("Corrupt PNG data " + path.string());
    std::vector<uint8_t> img(stride * (size_t)h);
    for (int y = 0; y < h; ++y) {
        const uint8_t f = raw[(stride + 1) * y];
        const uint8_t* src = &raw[(stride + 1) * y + 1];
        uint8_t* cur = &img[stride * y];
        const uint8_t* prev = y ? &img[stride * (y - 1)] : nullptr;
        for (size_t i = 0; i < stride; ++i) {
            const int a = i >= (size_t)ch ? cur[i - ch] : 0, b = prev ? prev[i] : 0;
            const int c = (prev && i >= (size_t)ch) ? prev[i - ch] : 0;
            int pred = 0;
            switch (f) {
            case 0: pred = 0; break;
            case 1: pred = a; break;
            case 2: pred = b; break;
            case 3: pred = (a + b) >> 1; break;
            case 4: {
                const int p = a + b - c, pa = std::abs(p - a), pb = std::abs(p - b), pc = std::abs(p - c);
                pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                break;
            }
            default: throw std::runtime_error("Bad PNG filter " + path.string());
            }
            cur[i] = (uint8_t)(src[i] + pred);
        }
    }
    Pgm p;
    p.width = w;
    p.height = h;
    p.data.resize((size_t)w * h);
    for (size_t i = 0; i < p.data.size(); ++i) {
        const uint8_t* px = &img[i * ch];
        if (ctype == 0 || ctype == 4) {
            p.data[i] = px[0];
        } else if (ctype == 3) {
            if ((size_t)px[0] * 3 + 2 >= plte.size()) throw std::runtime_error("PNG palette index " + path.string());
            const uint8_t* e = &plte[(size_t)px[0] * 3];
            p.data[i] = rgb_gray(e[0], e[1], e[2]);
        } else {
            p.data[i] = rgb_gray(px[0], px[1], px[2]);
        }
    }
    return p;
}

Pgm read_image(const fs::path& path) {
    std::string ext = path.extension().string();
    std::transform(ext.begin(), ext.end(), ext.begin(), ::tolower);
    return ext == ".png" ? read_png(path) : read_pgm(path);
}

void write_pgm(const fs::path& path, int w, int h, const uint8_t* data) {
    std::ofstream out(path, std::ios::binary);
    out << "P5\n" << w << ' ' << h << "\n255\n";
    out.write(reinterpret_cast<const char*>(data), (std::streamsize)w * h);
}

// Reference quantiser (harness/quantization_and_scaling/gpu_common_quant_v0_7.cu:479-495).
void quantize(std::vector<uint8_t>& img, int levels) {
    if (levels == 256) return;
    for (auto& v : img) {
        const int bin = ((int)v * levels) >> 8;
        v = (uint8_t)((bin * 255 + (levels - 1) / 2) / (levels - 1));
    }
}

double percentile(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = p * (double)(v.size() - 1);
    const size_t lo = (size_t)std::floor(pos), hi = (size_t)std::ceil(pos);
    const double f = pos - (double)lo;
    return v[lo] * (1.0 - f) + v[hi] * f;
}

uint64_t compare_masks(const std::vector<uint8_t>& a, const std::vector<uint8_t>& b) {
    if (a.size() != b.size()) throw std::runtime_error("Mask dimensions differ");
    uint64_t m = 0;
    for (size_t i = 0; i < a.size(); ++i) m += a[i] != b[i];
    return m;
}

std::string quote(const std::string& s) {
    std::string o = "\"";
    for (char c : s) {
        if (c == '"') o.push_back('"');
        o.push_back(c);
    }
    return o + "\"";
}

bool starts(const std::string& s, const char* p) { return s.rfind(p, 0) == 0; }

std::vector<int> int_list(const std::string& v) {
    std::vector<int> out;
    size_t b = 0;
    while (b < v.size()) {
        const size_t c = v.find(',', b);
        out.push_back(std::stoi(v.substr(b, c == std::string::npos ? std::string::npos : c - b)));
        if (c == std::string::npos) break;
        b = c + 1;
    }
    return out;
}

std::vector<fs::path> image_files(const fs::path& dir) {
    std::vector<fs::path> files;
    for (const auto& e : fs::directory_iterator(dir)) {
        std::string ext = e.path().extension().string();
        std::transform(ext.begin(), ext.end(), ext.begin(), ::tolower);
        if (e.is_regular_file() && (ext == ".pgm" || ext == ".png")) files.push_back(e.path());
    }
    std::sort(files.begin(), files.end());
    return files;
}

int convert_main(int argc, char** argv) {
    if (argc < 4 || !starts(argv[3], "--output=")) {
        std::cerr << "Usage: ss_harness --convert <image_dir> --output=DIR\n";
        return 2;
    }
    try {
        const fs::path out = std::string(argv[3]).substr(9);
        fs::create_directories(out);
        const std::vector<fs::path> files = image_files(argv[2]);
        for (const fs::path& f : files) {
            const Pgm p = read_image(f);
            write_pgm(out / (f.stem().string() + ".pgm"), p.width, p.height, p.data.data());
        }
        std::cout << "[SS-CONVERT-OK] images=" << files.size() << " output=" << out.string() << '\n';
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "[ERROR] " << e.what() << '\n';
        return 3;
    }
}

int seam_profile_main(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "Usage: ss_harness --seam-profile <image_dir> --output=FILE [--tiles=32,64,128]"
                     " [--levels=256,32,8,2] [--device=D]\n";
        return 2;
    }
    try {
        const fs::path image_dir = argv[2];
        fs::path output;
        std::vector<int> tiles{32, 64, 128}, levels{256, 32, 8, 2};
        int device = 0;
        for (int i = 3; i < argc; ++i) {
            const std::string a = argv[i];
            if (starts(a, "--output=")) output = a.substr(9);
            else if (starts(a, "--tiles=")) tiles = int_list(a.substr(8));
            else if (starts(a, "--levels=")) levels = int_list(a.substr(9));
            else if (starts(a, "--device=")) device = std::stoi(a.substr(9));
            else {
                std::cerr << "Unknown argument " << a << '\n';
                return 2;
            }
        }
        if (output.empty()) throw std::runtime_error("Missing --output file");
        for (int t : tiles)
            if (t != 16 && t != 32 && t != 64 && t != 128)
                throw std::runtime_error("Tile size must be 16, 32, 64 or 128");
        for (int l : levels)
            if (l < 2 || l > 256) throw std::runtime_error("Levels must be in [2,256]");
        SS_HIP_CHECK(hipSetDevice(device));
        if (output.has_parent_path()) fs::create_directories(output.parent_path());
        std::ofstream csv(output);
        if (!csv) throw std::runtime_error("Cannot create output CSV");
        csv << "image,width,height,megapixels,levels,tile,"
               "total_tile_fragments,seam_fragments,interior_fragments,"
               "seam_fragment_fraction,seam_pixels,seam_pixel_fraction,"
               "equal_forward_edges,equal_seam_edges,"
               "seam_edge_fraction,seam_fragments_per_mpix,"
               "equal_seam_edges_per_mpix,diagnostic_ms\n";
        const std::vector<fs::path> files = image_files(image_dir);
        for (const fs::path& f : files) {
            const Pgm src = read_image(f);
            const size_t px = src.data.size();
            const double mp = (double)px / 1e6;
            uint8_t* d_img = nullptr;
            void* d_ws = nullptr;
            const size_t ws_bytes = 256;   // the profiler's counters only
            SS_HIP_CHECK(hipMalloc(&d_img, px));
            SS_HIP_CHECK(hipMalloc(&d_ws, ws_bytes));
            for (int level : levels) {
                std::vector<uint8_t> q = src.data;
                quantize(q, level);
                SS_HIP_CHECK(hipMemcpy(d_img, q.data(), px, hipMemcpyHostToDevice));
                for (int t : tiles) {
                    const ss_config c{t, 256};
                    ss_seam_profile p{};
                    const auto t0 = std::chrono::high_resolution_clock::now();
                    sparse_seam::check(ss_seam_profile_u8(d_img, (size_t)src.width, src.width, src.height, &c,
                                                          d_ws, ws_bytes, nullptr, &p),
                                       "ss_seam_profile_u8");
                    const double ms =
                        std::chrono::duration<double, std::milli>(std::chrono::high_resolution_clock::now() - t0)
                            .count();
                    const double sff = p.fragments ? (double)p.seam_fragments / (double)p.fragments : 0.0;
                    const double spf = (double)p.seam_pixels / (double)px;
                    const double sef = p.equal_edges ? (double)p.seam_edges / (double)p.equal_edges : 0.0;
                    csv << quote(f.filename().string()) << ',' << src.width << ',' << src.height << ','
                        << std::fixed << std::setprecision(6) << mp << ',' << level << ',' << t << ','
                        << p.fragments << ',' << p.seam_fragments << ',' << p.fragments - p.seam_fragments << ','
                        << sff << ',' << p.seam_pixels << ',' << spf << ',' << p.equal_edges << ','
                        << p.seam_edges << ',' << sef << ',' << (double)p.seam_fragments / mp << ','
                        << (double)p.seam_edges / mp << ',' << ms << '\n';
                    csv.flush();
                    std::cout << "[SS-SEAM-PROFILE] " << f.filename().string() << " q=" << level << " tile=" << t
                              << " fragments=" << p.fragments << " seam_fragments=" << p.seam_fragments
                              << " seam_fragment_fraction=" << sff << " seam_pixel_fraction=" << spf
                              << " seam_edges=" << p.seam_edges << '\n';
                }
            }
            SS_HIP_CHECK(hipFree(d_img));
            SS_HIP_CHECK(hipFree(d_ws));
        }
        std::cout << "[SS-SEAM-PROFILE-OK] images=" << files.size() << " output=" << output.string() << '\n';
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "[ERROR] " << e.what() << '\n';
        return 3;
    }
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 2 && std::string(argv[1]) == "--seam-profile") return seam_profile_main(argc, argv);
    if (argc >= 2 && std::string(argv[1]) == "--convert") return convert_main(argc, argv);
    if (argc < 3) {
        std::cerr << "Usage: ss_harness <image_dir> --output=DIR [--expected-max=DIR --expected-min=DIR]"
                     " [--warmup=5] [--reps=30] [--tile=T --threads=N | --autotune] [--levels=L]"
                     " [--device=D] [--skip-expected]\n";
        return 2;
    }
    try {
        const fs::path image_dir = argv[1];
        fs::path out_dir, exp_max, exp_min;
        int warmup = 5, reps = 30, device = 0, levels = 256;
        bool skip_expected = false, autotune = false;
        ss_config cfg{};
        sparse_seam::check(ss_get_default_config(&cfg), "ss_get_default_config");
        for (int i = 2; i < argc; ++i) {
            const std::string a = argv[i];
            if (starts(a, "--output=")) out_dir = a.substr(9);
            else if (starts(a, "--expected-max=")) exp_max = a.substr(15);
            else if (starts(a, "--expected-min=")) exp_min = a.substr(15);
            else if (starts(a, "--warmup=")) warmup = std::stoi(a.substr(9));
            else if (starts(a, "--reps=")) reps = std::stoi(a.substr(7));
            else if (starts(a, "--tile=")) cfg.tile = std::stoi(a.substr(7));
            else if (starts(a, "--threads=")) cfg.threads = std::stoi(a.substr(10));
            else if (starts(a, "--levels=")) levels = std::stoi(a.substr(9));
            else if (starts(a, "--device=")) device = std::stoi(a.substr(9));
            else if (a == "--skip-expected") skip_expected = true;
            else if (a == "--autotune") autotune = true;
            else {
                std::cerr << "Unknown argument " << a << '\n';
                return 2;
            }
        }
        if (out_dir.empty()) throw std::runtime_error("Missing --output directory");
        if (!skip_expected && (exp_max.empty() || exp_min.empty()))
            throw std::runtime_error("Missing --expected-max/--expected-min (or --skip-expected)");
        if (levels < 2 || levels > 256) throw std::runtime_error("--levels must be in 2..256");
        if (reps < 1 || warmup < 0) throw std::runtime_error("--reps must be >= 1 and --warmup >= 0");
        SS_HIP_CHECK(hipSetDevice(device));
        hipDeviceProp_t prop{};
        SS_HIP_CHECK(hipGetDeviceProperties(&prop, device));

        const std::vector<fs::path> files = image_files(image_dir);

        std::vector<Pgm> images;
        for (const auto& f : files) {
            images.push_back(read_image(f));
            quantize(images.back().data, levels);
        }

        if (autotune && !images.empty()) {
            // time the candidates on (up to) the first 8 images
            std::vector<sparse_seam::DeviceBuffers> tb(std::min<size_t>(8, images.size()));
            std::vector<ss_image_desc> d(tb.size());
            for (size_t i = 0; i < tb.size(); ++i) {
                sparse_seam::allocate_buffers(tb[i], images[i].width, images[i].height, &cfg);
                SS_HIP_CHECK(hipMemcpy(tb[i].image, images[i].data.data(), images[i].data.size(),
                                       hipMemcpyHostToDevice));
                d[i] = ss_image_desc{tb[i].image, tb[i].maxima, tb[i].minima, (uint64_t)images[i].width,
                                     (uint64_t)images[i].width, images[i].width, images[i].height};
            }
            double rate = 0;
            sparse_seam::check(ss_autotune(d.data(), (int)d.size(), 5, nullptr, &cfg, &rate), "ss_autotune");
            for (auto& b : tb) sparse_seam::free_buffers(b);
            std::cout << "[SS-AUTOTUNE] tile=" << cfg.tile << " threads=" << cfg.threads
                      << " Mpix_s=" << rate / 1e6 << '\n';
        }

        fs::create_directories(out_dir / "actual_max");
        fs::create_directories(out_dir / "actual_min");
        std::ofstream csv(out_dir / "ss_harness.csv");
        csv << "algorithm,image,width,height,megapixels,warmup,repetitions,"
               "block,total_allocated_bytes,total_allocated_Bpp,"
               "auxiliary_bytes,auxiliary_Bpp,"
               "median_compute_ms,p95_compute_ms,compute_mpix_s,"
               "median_h2h_ms,p95_h2h_ms,h2h_mpix_s,"
               "median_iterations,p95_iterations,"
               "oracle_max_mismatches,oracle_min_mismatches,"
               "cross_max_mismatches,cross_min_mismatches,"
               "tile,active_fragments,dirty_tiles,dsu_retries\n";
        // per-image validation rows with the columns of the reference's
        // snapshot harness (snapshot_sparse_seam_cuda_v1_0.cu:931-933,1000-1014);
        // auxiliary bytes are the library's workspace (the masks and the image
        // are the caller's)
        std::ofstream vcsv(out_dir / "validation.csv");
        vcsv << "image,width,height,tiles,fragment_capacity,active_fragments,"
                "auxiliary_bytes,auxiliary_bytes_per_pixel,elapsed_ms,"
                "oracle_checked,max_mismatches,min_mismatches\n";
        {
            std::ofstream env(out_dir / "environment.txt");
            env << "harness=ss_harness\n"
                << "device_name=" << prop.name << '\n'
                << "gcn_arch=" << prop.gcnArchName << '\n'
                << "compute_units=" << prop.multiProcessorCount << '\n'
                << "global_memory_bytes=" << prop.totalGlobalMem << '\n'
                << "tile=" << cfg.tile << '\n'
                << "threads_per_tile=" << cfg.threads << '\n'
                << "warmup=" << warmup << '\n'
                << "repetitions=" << reps << '\n'
                << "levels=" << levels << '\n'
                << "timing_compute=preallocated,resident_input,task_complete\n"
                << "timing_h2h=reused_allocation,H2D+compute+D2H_two_masks\n"
                << "auxiliary_definition=workspace_bytes(all device buffers except input and two masks)\n"
                << "cross_route=Global-UF (ss_global_uf_u8) on every image, masks compared byte for byte\n"
                << "iterations=value returned by run_sparse_resident per timed run\n";
        }

        uint64_t total_mismatch = 0, total_cross = 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        SS_HIP_CHECK(hipEventCreate(&e0));
        SS_HIP_CHECK(hipEventCreate(&e1));
        for (size_t k = 0; k < files.size(); ++k) {
            const Pgm& img = images[k];
            const size_t px = img.data.size();
            sparse_seam::DeviceBuffers b;
            sparse_seam::allocate_buffers(b, img.width, img.height, &cfg);
            std::vector<uint8_t> mx(px), mn(px);
            SS_HIP_CHECK(hipMemcpy(b.image, img.data.data(), px, hipMemcpyHostToDevice));
            for (int i = 0; i < warmup; ++i) {
                sparse_seam::run_sparse_resident(b);
                SS_HIP_CHECK(hipDeviceSynchronize());
            }
            std::vector<double> compute, h2h, iterations;
            for (int i = 0; i < reps; ++i) {
                SS_HIP_CHECK(hipEventRecord(e0, nullptr));
                iterations.push_back((double)sparse_seam::run_sparse_resident(b));
                SS_HIP_CHECK(hipEventRecord(e1, nullptr));
                SS_HIP_CHECK(hipEventSynchronize(e1));
                float ms = 0.f;
                SS_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
                compute.push_back(ms);
            }
            for (int i = 0; i < reps; ++i) {
                const auto t0 = std::chrono::high_resolution_clock::now();
                SS_HIP_CHECK(hipMemcpy(b.image, img.data.data(), px, hipMemcpyHostToDevice));
                sparse_seam::run_sparse_resident(b);
                SS_HIP_CHECK(hipMemcpy(mx.data(), b.maxima, px, hipMemcpyDeviceToHost));
                SS_HIP_CHECK(hipMemcpy(mn.data(), b.minima, px, hipMemcpyDeviceToHost));
                const auto t1 = std::chrono::high_resolution_clock::now();
                h2h.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
            }
            const ss_stats st = sparse_seam::read_stats(b);
            // cross-route check against the dense Global-UF route on the same
            // resident image (gpu_common_v0_6.cu:659-669)
            uint64_t cross_max = 0, cross_min = 0;
            {
                size_t gws = 0;
                sparse_seam::check(ss_global_uf_workspace_size(img.width, img.height, &gws),
                                   "ss_global_uf_workspace_size");
                void* g = nullptr;
                SS_HIP_CHECK(hipMalloc(&g, gws));
                std::vector<uint8_t> gx(px), gn(px);
                sparse_seam::run_global_uf(b, g, gws);
                SS_HIP_CHECK(hipMemcpy(gx.data(), b.maxima, px, hipMemcpyDeviceToHost));
                SS_HIP_CHECK(hipMemcpy(gn.data(), b.minima, px, hipMemcpyDeviceToHost));
                SS_HIP_CHECK(hipFree(g));
                cross_max = compare_masks(gx, mx);
                cross_min = compare_masks(gn, mn);
            }
            total_cross += cross_max + cross_min;
            uint64_t mm_max = 0, mm_min = 0;
            const std::string stem = files[k].stem().string();
            if (!skip_expected) {
                const Pgm em = read_pgm(exp_max / (stem + ".pgm"));
                const Pgm en = read_pgm(exp_min / (stem + ".pgm"));
                mm_max = compare_masks(mx, em.data);
                mm_min = compare_masks(mn, en.data);
                write_pgm(out_dir / "actual_max" / (stem + ".pgm"), img.width, img.height, mx.data());
                write_pgm(out_dir / "actual_min" / (stem + ".pgm"), img.width, img.height, mn.data());
            }
            total_mismatch += mm_max + mm_min;
            const double mp = (double)px / 1e6;
            const double med_c = percentile(compute, 0.5), med_h = percentile(h2h, 0.5);
            const uint64_t total_bytes = (uint64_t)px * 3 + b.workspace_bytes;
            csv << "Sparse-Seam-MI355X," << quote(files[k].filename().string()) << ',' << img.width << ','
                << img.height << ',' << std::fixed << std::setprecision(6) << mp << ',' << warmup << ','
                << reps << ',' << cfg.threads << ',' << total_bytes << ',' << (double)total_bytes / px << ','
                << b.workspace_bytes << ',' << (double)b.workspace_bytes / px << ',' << med_c << ','
                << percentile(compute, 0.95) << ',' << mp / (med_c * 1e-3) << ',' << med_h << ','
                << percentile(h2h, 0.95) << ',' << mp / (med_h * 1e-3) << ',' << percentile(iterations, 0.5) << ','
                << percentile(iterations, 0.95) << ',' << mm_max << ',' << mm_min << ',' << cross_max << ','
                << cross_min << ',' << cfg.tile << ',' << st.fragments << ',' << st.dirty_tiles << ',' << st.dsu_retries
                << '\n';
            csv.flush();
            const uint64_t tiles = (uint64_t)((img.width + cfg.tile - 1) / cfg.tile) *
                                   (uint64_t)((img.height + cfg.tile - 1) / cfg.tile);
            vcsv << quote(files[k].filename().string()) << ',' << img.width << ',' << img.height << ',' << tiles
                 << ',' << tiles * (uint64_t)(4 * cfg.tile - 4) << ',' << st.fragments << ','
                 << b.workspace_bytes << ',' << std::fixed << std::setprecision(6)
                 << (double)b.workspace_bytes / px << ',' << med_c << ',' << (skip_expected ? 0 : 1) << ','
                 << mm_max << ',' << mm_min << '\n';
            vcsv.flush();
            std::cout << "[SS-HARNESS] " << files[k].filename().string() << ' ' << img.width << 'x'
                      << img.height << " compute_ms=" << med_c << " h2h_ms=" << med_h
                      << " fragments=" << st.fragments << " max_mismatch=" << mm_max
                      << " min_mismatch=" << mm_min << " cross_mismatch=" << cross_max + cross_min << '\n';
            sparse_seam::free_buffers(b);
        }
        SS_HIP_CHECK(hipEventDestroy(e0));
        SS_HIP_CHECK(hipEventDestroy(e1));
        std::cout << (skip_expected ? "[SS-HARNESS-MEASUREMENT-OK]" : "[SS-HARNESS-VALIDATION-OK]")
                  << " images=" << files.size() << " mismatches=" << total_mismatch
                  << " cross_mismatches=" << total_cross << " output=" << out_dir.string() << '\n';
        return total_mismatch == 0 && total_cross == 0 ? 0 : 1;
    } catch (const std::exception& e) {
        std::cerr << "[ERROR] " << e.what() << '\n';
        return 3;
    }
}
