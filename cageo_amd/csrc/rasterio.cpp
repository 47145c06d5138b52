// Raster header/binary IO and a baseline grayscale JPEG writer.
#include "rasterio.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace cageo_host {

void loadHeaderData(const std::string& fileName, DemInfo* info) {
    std::ifstream f(fileName.c_str());
    if (!f.is_open()) throw std::runtime_error("Unable to open file \"" + fileName + "\"");
    std::string tag;
    while (f >> tag) {
        bool ok = true;
        if (tag == "ncols") ok = static_cast<bool>(f >> info->fileCols);
        else if (tag == "nrows") ok = static_cast<bool>(f >> info->fileRows);
        else if (tag == "xllcorner") ok = static_cast<bool>(f >> info->xllCorner);
        else if (tag == "yllcorner") ok = static_cast<bool>(f >> info->yllCorner);
        else if (tag == "cellsize") ok = static_cast<bool>(f >> info->cellSize);
        else if (tag == "NODATA_value") ok = static_cast<bool>(f >> info->noDataValue);
        else if (tag == "byteorder" || tag == "BYTEORDER") {
            std::string v;
            ok = static_cast<bool>(f >> v);
            if (v == "MSBFIRST") info->byteOrder = MSBFIRST;
            else if (v == "LSBFIRST") info->byteOrder = LSBFIRST;
            else if (ok) throw std::runtime_error("Unrecognized byte order value: " + v);
        } else {
            throw std::runtime_error("Unrecognized tag value: " + tag);
        }
        if (!ok) throw std::runtime_error("Bad value for tag " + tag + " in " + fileName);
    }
}

void writeHeader(const std::string& fileName, const DemInfo& info) {
    std::ofstream f(fileName.c_str());
    if (!f.is_open()) throw std::runtime_error("Unable to write \"" + fileName + "\"");
    f << "ncols " << info.fileCols << "\n"
      << "nrows " << info.fileRows << "\n"
      << "xllcorner " << info.xllCorner << "\n"
      << "yllcorner " << info.yllCorner << "\n"
      << "cellsize " << info.cellSize << "\n"
      << "NODATA_value " << info.noDataValue << "\n"
      << "byteorder " << (info.byteOrder == MSBFIRST ? "MSBFIRST" : "LSBFIRST") << "\n";
}

void readRaster(const std::string& fileName, float* data, size_t count) {
    std::ifstream f(fileName.c_str(), std::ios::binary);
    if (!f.is_open()) throw std::runtime_error("Unable to open file \"" + fileName + "\"");
    f.read(reinterpret_cast<char*>(data), (std::streamsize)(count * sizeof(float)));
    if ((size_t)f.gcount() != count * sizeof(float))
        throw std::runtime_error("File \"" + fileName + "\" holds fewer than " +
                                 std::to_string(count) + " float cells");
}

template <typename T>
void writeRaster(const std::string& fileName, const T* data, size_t count) {
    std::ofstream f(fileName.c_str(), std::ios::binary | std::ios::trunc);
    if (!f.is_open()) throw std::runtime_error("Unable to write \"" + fileName + "\"");
    f.write(reinterpret_cast<const char*>(data), (std::streamsize)(count * sizeof(T)));
    if (!f) throw std::runtime_error("Write failed for \"" + fileName + "\"");
}
template void writeRaster<float>(const std::string&, const float*, size_t);
template void writeRaster<int>(const std::string&, const int*, size_t);
template void writeRaster<unsigned char>(const std::string&, const unsigned char*, size_t);

// ---------------------------------------------------------------------------
// Baseline JPEG, one 8-bit component, standard tables of ITU-T T.81 Annex K.

namespace {

const int kZigzag[64] = {0,  1,  8,  16, 9,  2,  3,  10, 17, 24, 32, 25, 18, 11, 4,  5,
                         12, 19, 26, 33, 40, 48, 41, 34, 27, 20, 13, 6,  7,  14, 21, 28,
                         35, 42, 49, 56, 57, 50, 43, 36, 29, 22, 15, 23, 30, 37, 44, 51,
                         58, 59, 52, 45, 38, 31, 39, 46, 53, 60, 61, 54, 47, 55, 62, 63};

// Table K.1, natural (row-major) order.
const int kLumaQ[64] = {16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
                        14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
                        18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
                        49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};

// Tables K.3 (DC) and K.5 (AC), luminance.
const uint8_t kDcBits[16] = {0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0};
const uint8_t kDcVals[12] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11};
const uint8_t kAcBits[16] = {0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d};
const uint8_t kAcVals[162] = {
    0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51, 0x61,
    0x07, 0x22, 0x71, 0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1, 0x15, 0x52,
    0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72, 0x82, 0x09, 0x0a, 0x16, 0x17, 0x18, 0x19, 0x1a, 0x25,
    0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37, 0x38, 0x39, 0x3a, 0x43, 0x44, 0x45,
    0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57, 0x58, 0x59, 0x5a, 0x63, 0x64,
    0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x83,
    0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99,
    0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6,
    0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3,
    0xd4, 0xd5, 0xd6, 0xd7, 0xd8, 0xd9, 0xda, 0xe1, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8,
    0xe9, 0xea, 0xf1, 0xf2, 0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa};

struct Huff {
    uint16_t code[256];
    uint8_t len[256];
    // Canonical codes from the (bits, vals) lists, T.81 Annex C.
    Huff(const uint8_t* bits, const uint8_t* vals) {
        for (int i = 0; i < 256; ++i) len[i] = 0;
        int c = 0, k = 0;
        for (int l = 1; l <= 16; ++l) {
            for (int i = 0; i < bits[l - 1]; ++i, ++k) {
                code[vals[k]] = (uint16_t)c++;
                len[vals[k]] = (uint8_t)l;
            }
            c <<= 1;
        }
    }
};

class BitWriter {
public:
    explicit BitWriter(std::vector<uint8_t>* out) : out_(out) {}
    void put(uint32_t bits, int n) {
        acc_ = (acc_ << n) | (bits & ((1u << n) - 1));
        nacc_ += n;
        while (nacc_ >= 8) {
            const uint8_t b = (uint8_t)(acc_ >> (nacc_ - 8));
            out_->push_back(b);
            if (b == 0xff) out_->push_back(0);  // byte stuffing
            nacc_ -= 8;
        }
        acc_ &= (1u << nacc_) - 1;
    }
    void flush() {
        if (nacc_ > 0) put(0x7f, 8 - nacc_);  // pad with 1-bits
    }

private:
    std::vector<uint8_t>* out_;
    uint32_t acc_ = 0;
    int nacc_ = 0;
};

int category(int v) {
    int a = v < 0 ? -v : v, n = 0;
    while (a) {
        ++n;
        a >>= 1;
    }
    return n;
}

void put16(std::vector<uint8_t>& o, int v) {
    o.push_back((uint8_t)(v >> 8));
    o.push_back((uint8_t)v);
}

}  // namespace

void writeJpegGray(const std::string& fileName, const unsigned char* px, int W, int H, int quality) {
    if (W <= 0 || H <= 0 || W > 65535 || H > 65535)
        throw std::runtime_error("JPEG: image size out of range");
    quality = quality < 1 ? 1 : (quality > 100 ? 100 : quality);
    const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;
    int q[64];
    for (int i = 0; i < 64; ++i) {
        int v = (kLumaQ[i] * scale + 50) / 100;
        q[i] = v < 1 ? 1 : (v > 255 ? 255 : v);
    }
    double cosT[8][8];
    for (int x = 0; x < 8; ++x)
        for (int u = 0; u < 8; ++u)
            cosT[x][u] = std::cos((2 * x + 1) * u * M_PI / 16.0) * (u == 0 ? std::sqrt(0.5) : 1.0);

    std::vector<uint8_t> o;
    o.reserve((size_t)W * H / 4 + 1024);
    o.push_back(0xff); o.push_back(0xd8);  // SOI
    o.push_back(0xff); o.push_back(0xe0);  // APP0 JFIF
    put16(o, 16);
    const char jfif[5] = {'J', 'F', 'I', 'F', 0};
    o.insert(o.end(), jfif, jfif + 5);
    o.push_back(1); o.push_back(1); o.push_back(0);
    put16(o, 1); put16(o, 1);
    o.push_back(0); o.push_back(0);
    o.push_back(0xff); o.push_back(0xdb);  // DQT
    put16(o, 67);
    o.push_back(0);
    for (int k = 0; k < 64; ++k) o.push_back((uint8_t)q[kZigzag[k]]);
    o.push_back(0xff); o.push_back(0xc0);  // SOF0
    put16(o, 11);
    o.push_back(8);
    put16(o, H); put16(o, W);
    o.push_back(1); o.push_back(1); o.push_back(0x11); o.push_back(0);
    o.push_back(0xff); o.push_back(0xc4);  // DHT
    put16(o, 2 + 17 + 12 + 17 + 162);
    o.push_back(0x00);
    o.insert(o.end(), kDcBits, kDcBits + 16);
    o.insert(o.end(), kDcVals, kDcVals + 12);
    o.push_back(0x10);
    o.insert(o.end(), kAcBits, kAcBits + 16);
    o.insert(o.end(), kAcVals, kAcVals + 162);
    o.push_back(0xff); o.push_back(0xda);  // SOS
    put16(o, 8);
    o.push_back(1); o.push_back(1); o.push_back(0x00);
    o.push_back(0); o.push_back(63); o.push_back(0);

    const Huff dc(kDcBits, kDcVals), ac(kAcBits, kAcVals);
    BitWriter bw(&o);
    int prevDc = 0;
    double blk[64], tmp[64];
    for (int by = 0; by < H; by += 8)
        for (int bx = 0; bx < W; bx += 8) {
            for (int y = 0; y < 8; ++y)
                for (int x = 0; x < 8; ++x) {
                    const int sx = bx + x < W ? bx + x : W - 1;  // replicate the edge
                    const int sy = by + y < H ? by + y : H - 1;
                    blk[y * 8 + x] = (double)px[(size_t)sy * W + sx] - 128.0;
                }
            // separable forward DCT-II with the T.81 normalisation (1/4 C(u) C(v))
            for (int y = 0; y < 8; ++y)
                for (int u = 0; u < 8; ++u) {
                    double s = 0;
                    for (int x = 0; x < 8; ++x) s += blk[y * 8 + x] * cosT[x][u];
                    tmp[y * 8 + u] = s;
                }
            int coef[64];
            for (int v = 0; v < 8; ++v)
                for (int u = 0; u < 8; ++u) {
                    double s = 0;
                    for (int y = 0; y < 8; ++y) s += tmp[y * 8 + u] * cosT[y][v];
                    s *= 0.25;
                    long c = std::lround(s / q[v * 8 + u]);
                    // baseline Huffman categories stop at 11 (DC) and 10 (AC)
                    coef[v * 8 + u] = (int)(c > 1023 ? 1023 : (c < -1023 ? -1023 : c));
                }
            const int diff = coef[0] - prevDc;
            prevDc = coef[0];
            int cat = category(diff);
            bw.put(dc.code[cat], dc.len[cat]);
            if (cat) bw.put(diff < 0 ? (uint32_t)(diff - 1) : (uint32_t)diff, cat);
            int run = 0;
            for (int k = 1; k < 64; ++k) {
                const int v = coef[kZigzag[k]];
                if (v == 0) {
                    ++run;
                    continue;
                }
                while (run > 15) {
                    bw.put(ac.code[0xf0], ac.len[0xf0]);
                    run -= 16;
                }
                cat = category(v);
                const int sym = (run << 4) | cat;
                bw.put(ac.code[sym], ac.len[sym]);
                bw.put(v < 0 ? (uint32_t)(v - 1) : (uint32_t)v, cat);
                run = 0;
            }
            if (run) bw.put(ac.code[0x00], ac.len[0x00]);
        }
    bw.flush();
    o.push_back(0xff); o.push_back(0xd9);  // EOI
    std::ofstream f(fileName.c_str(), std::ios::binary | std::ios::trunc);
    if (!f.is_open()) throw std::runtime_error("Unable to write \"" + fileName + "\"");
    f.write(reinterpret_cast<const char*>(o.data()), (std::streamsize)o.size());
}

}  // namespace cageo_host

// ---------------------------------------------------------------------------
// C ABI of the IO helpers (include/cageo_hip.h).
#include "cageo_hip.h"

extern "C" {

int cageo_read_header(const char* path, int* cols, int* rows, double* xll, double* yll,
                      float* cell_size, float* nodata) {
    if (!path || !cols || !rows) return CAGEO_ERR_ARG;
    cageo_host::DemInfo info;
    try {
        cageo_host::loadHeaderData(path, &info);
    } catch (const std::exception&) {
        return CAGEO_ERR_ARG;
    }
    *cols = info.fileCols;
    *rows = info.fileRows;
    if (xll) *xll = info.xllCorner;
    if (yll) *yll = info.yllCorner;
    if (cell_size) *cell_size = info.cellSize;
    if (nodata) *nodata = info.noDataValue;
    return CAGEO_OK;
}

int cageo_write_jpeg_gray(const char* path, const uint8_t* pixels, int width, int height,
                          int quality) {
    if (!path || !pixels) return CAGEO_ERR_ARG;
    try {
        cageo_host::writeJpegGray(path, pixels, width, height, quality);
    } catch (const std::exception&) {
        return CAGEO_ERR_ARG;
    }
    return CAGEO_OK;
}

}  // extern "C"
