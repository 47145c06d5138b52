/*
 * Raster input/output of the drainage program.
 *
 *  - loadHeaderData: the ESRI-style .hdr text header (utils.cpp:58-112):
 *    whitespace-separated tag/value pairs ncols, nrows, xllcorner,
 *    yllcorner, cellsize, NODATA_value, byteorder|BYTEORDER.  An unknown tag
 *    or byte order is an error.  The byte order is recorded but, as in the
 *    reference, the raw data are read as they are (native little-endian).
 *  - readRaster / writeRaster: raw float32 .bin files of W*H cells
 *    (SystemUtils.h:53-86).
 *  - writeHeader: the same header format, for the synthetic inputs.
 *
 * Failures throw std::runtime_error; the CLI turns them into the reference's
 * message-and-exit behaviour.
 */
#ifndef CAGEO_RASTERIO_HPP_
#define CAGEO_RASTERIO_HPP_

#include <string>
#include <vector>

namespace cageo_host {

enum ByteOrder { MSBFIRST, LSBFIRST, VMS_FFLOAT };

/* DemInfo.h:21-29. */
struct DemInfo {
    int fileCols = 0;
    int fileRows = 0;
    double xllCorner = 0.0;
    double yllCorner = 0.0;
    float cellSize = 0.0f;
    float noDataValue = 0.0f;
    ByteOrder byteOrder = LSBFIRST;
};

void loadHeaderData(const std::string& fileName, DemInfo* info);
void writeHeader(const std::string& fileName, const DemInfo& info);

/* Reads exactly count values; a short file is an error. */
void readRaster(const std::string& fileName, float* data, size_t count);
template <typename T>
void writeRaster(const std::string& fileName, const T* data, size_t count);

/* Baseline grayscale JPEG (ITU-T T.81, Huffman, 8-bit) of a W x H byte
 * image, the output format of main.cpp:254-266.  quality 1..100 scales the
 * T.81 Annex K luminance table as the IJG convention does. */
void writeJpegGray(const std::string& fileName, const unsigned char* pixels, int width, int height,
                   int quality = 90);

}  // namespace cageo_host

#endif  // CAGEO_RASTERIO_HPP_
