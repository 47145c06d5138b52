/*
 * drainage: Monte Carlo drainage-basin border frequencies on one MI355X.
 *
 * The command line and the flow of the reference's main.cpp (options
 * :329-372 and :383-416, set-up :76-156, Monte Carlo loop :180-283, timing
 * report :285-313), with these additions:
 *   --seed N              Philox key (the reference seeds with time()+i,
 *                         :189-192); realization i uses counter word i, so a
 *                         run is reproducible and independent of chunking.
 *   --staged              run every stage through its plugin class and
 *                         factory, one blocking call per stage as main.cpp
 *                         does; the default runs the same stages fused on
 *                         the device (cageo_run_realizations), with the same
 *                         result.
 *   --timing              per-stage timing report.
 *   --checkpoint F, --checkpoint-every K, --resume F
 *                         write/read a resume point: the counter raster and
 *                         the next realization index.
 *   --synth W,H,SEED      write a synthetic DEM and stream network (headers
 *                         and .bin) under --dem-path before running.
 * --cpu selects the reference's own CPU classes, which are not part of this
 * program; it is refused with a message.
 */
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "CellGrid.h"
#include "rasterio.hpp"
#include "Stages.h"
#include "Timer.h"

using namespace cageo_host;

namespace {

struct Options {
    std::string demName = "dem10", streamsName = "streams";
    std::string demPath = "DEM", streamsPath = "DEM";
    FilterModel filterModel = GAUSSIAN;
    float practicalRange = 60.0f;
    float errorMean = 0.0f, errorStd = 1.0f;
    int iterations = 10;
    bool useGPU = true, useCPU = false;
    bool jpeg = true;
    enum { OUT_NONE, OUT_END, OUT_FRAMES } outputOption = OUT_NONE;
    std::string outputName = "result";
    int bevelRadius = 2;  // G::bevelRadius, GlobalParameters.cpp:46
    // additions
    bool seedGiven = false;
    uint32_t seed = 0;
    bool staged = false, timing = false;
    std::string checkpoint, resume;
    int checkpointEvery = 0;
    int synthW = 0, synthH = 0;
    uint64_t synthSeed = 0;
    int accThreshold = 2000;
    bool fastConvolution = false;
    std::string surrogate;
    int surrogateCheck = 8;
    double surrogateL1 = 0.02;
    double surrogateRmsRatio = 2.46;
    int surrogateWindow = 0;
    int surrogateRefine = 200;
    std::string batchFile;
    std::string batchMode = "graph";
    bool batchTune = false;
};

const char* kHelp =
    "Options:\n"
    "  --help                              Produce help message\n"
    "  --dem arg (=dem10)                  Name of the input DEM\n"
    "  --stream arg (=streams)             Name of the input stream network\n"
    "  --dem-path arg (=DEM)               Path to the input DEM\n"
    "  --stream-path arg (=DEM)            Path to the input stream network (as in the\n"
    "                                      reference, the stream files are read from\n"
    "                                      --dem-path)\n"
    "  --filter-type arg                   Spatial autocorrelation model for the random\n"
    "                                      field ('Gaussian' or 'Exponential')\n"
    "  --practical-range arg (=60)         Practical range of the autocorrelation model\n"
    "  --error-mean arg (=0)               Mean of the random field\n"
    "  --error-standard-deviation arg (=1) Standard deviation of the random field\n"
    "  -i [ --iterations ] arg (=10)       Number of Monte Carlo iterations\n"
    "  --gpu [=arg(=1)] (=1)               Run algorithm on GPU\n"
    "  --cpu [=arg(=1)] (=0)               Run algorithm on CPU (not in this build)\n"
    "  -o [ --output-file ] arg            Write result into file (*.bin or *.jpg)\n"
    "  --output-option arg                 end | frames\n"
    "  --seed arg                          Random field seed (default: time)\n"
    "  --staged                            Run the stage classes one by one\n"
    "  --timing                            Print the per-stage timing report\n"
    "  --checkpoint arg                    Resume-point file to write\n"
    "  --checkpoint-every arg (=0)         Write the resume point every K iterations\n"
    "  --resume arg                        Continue from a resume-point file\n"
    "  --synth W,H,SEED                    Generate synthetic inputs under --dem-path\n"
    "  --fast-convolution                  Exponential convolution through FFTs of the\n"
    "                                      grid's torus (power-of-two sides; else the\n"
    "                                      filter's low-rank separable form): within 1e-5\n"
    "                                      of the exact order, not bit-exact\n"
    "  --surrogate arg                     Border-probability network weights (.bin):\n"
    "                                      predict, check with a short Monte Carlo, and\n"
    "                                      write probability x iterations; the full Monte\n"
    "                                      Carlo runs if the check fails\n"
    "  --surrogate-check arg (=8)          Realizations of the check (>= 2)\n"
    "  --surrogate-l1 arg (=0.02)          Mean absolute error the map may have: the\n"
    "                                      check accepts a mean signed error up to half\n"
    "                                      of it and an RMS error up to 1.2 x ratio x it\n"
    "  --surrogate-rms-ratio arg (=2.46)   RMS / mean absolute error of the network's\n"
    "                                      errors (measured for the packaged weights)\n"
    "  --surrogate-window arg (=0)         Check windows of N x N cells; failing windows\n"
    "                                      take a --surrogate-refine Monte Carlo (0: one\n"
    "                                      window, the full Monte Carlo on failure)\n"
    "  --surrogate-refine arg (=200)       Realizations that refine failing windows\n"
    "  --batch arg                         File of tiles, one per line: DEM name, stream\n"
    "                                      name [, error std, practical range [, filter\n"
    "                                      type]] (rasters under --dem-path; missing\n"
    "                                      fields from the options above); all tiles run\n"
    "                                      --iterations realizations together, tile t\n"
    "                                      seeded seed + t; output <name>_<DEM>.bin|.jpg\n"
    "  --batch-mode arg (=graph)           graph | stream | per-tile\n"
    "  --batch-tune                        Autotune the launch settings first\n";

[[noreturn]] void fail(const std::string& msg) {
    std::cout << msg << std::endl;
    std::exit(EXIT_FAILURE);
}

bool parseBool(const std::string& v) {
    if (v == "1" || v == "true" || v == "on" || v == "yes") return true;
    if (v == "0" || v == "false" || v == "off" || v == "no") return false;
    fail("Bad boolean value: " + v);
}

template <typename T>
T parseNum(const std::string& name, const std::string& v) {
    std::istringstream is(v);
    T out;
    if (!(is >> out) || !is.eof()) fail("Bad value for --" + name + ": " + v);
    return out;
}

std::string fileStem(const std::string& path, std::string* ext) {
    const size_t slash = path.find_last_of('/');
    std::string base = slash == std::string::npos ? path : path.substr(slash + 1);
    const size_t dot = base.find_last_of('.');
    if (dot == std::string::npos || dot == 0) {
        *ext = "";
        return base;
    }
    *ext = base.substr(dot);
    return base.substr(0, dot);
}

Options parseArgs(int argc, char** argv) {
    Options o;
    std::string outputFile, outputOption, filterType;
    bool haveOutputFile = false, haveOutputOption = false, haveFilter = false;
    const std::map<std::string, std::string> shortNames = {{"i", "iterations"}, {"o", "output-file"}};
    const std::map<std::string, bool> implicitBool = {{"gpu", true}, {"cpu", true}, {"staged", true},
                                                      {"timing", true}, {"help", true},
                                                      {"fast-convolution", true}, {"batch-tune", true}};
    for (int k = 1; k < argc; ++k) {
        std::string a = argv[k], name, value;
        bool hasValue = false;
        if (a.rfind("--", 0) == 0) {
            name = a.substr(2);
            const size_t eq = name.find('=');
            if (eq != std::string::npos) {
                value = name.substr(eq + 1);
                name = name.substr(0, eq);
                hasValue = true;
            }
        } else if (a.size() >= 2 && a[0] == '-') {
            auto it = shortNames.find(a.substr(1, 1));
            if (it == shortNames.end()) fail("unrecognised option '" + a + "'");
            name = it->second;
            if (a.size() > 2) {
                value = a.substr(2);
                hasValue = true;
            }
        } else {
            fail("unexpected argument '" + a + "'");
        }
        static const char* const known[] = {
            "help", "dem", "stream", "dem-path", "stream-path", "filter-type", "practical-range",
            "error-mean", "error-standard-deviation", "iterations", "gpu", "cpu", "staged", "timing",
            "output-file", "output-option", "seed", "checkpoint", "checkpoint-every", "resume", "synth",
            "fast-convolution", "surrogate", "surrogate-check", "surrogate-l1",
            "surrogate-rms-ratio", "surrogate-window", "surrogate-refine", "batch", "batch-mode",
            "batch-tune"};
        bool isKnown = false;
        for (const char* kname : known) isKnown = isKnown || name == kname;
        if (!isKnown) fail("unrecognised option '--" + name + "'");
        if (!hasValue && !implicitBool.count(name)) {
            if (k + 1 >= argc) fail("the required argument for option '--" + name + "' is missing");
            value = argv[++k];
            hasValue = true;
        }
        if (name == "help") {
            std::cout << kHelp << "\n";
            std::exit(EXIT_SUCCESS);
        } else if (name == "dem") o.demName = value;
        else if (name == "stream") o.streamsName = value;
        else if (name == "dem-path") o.demPath = value;
        else if (name == "stream-path") o.streamsPath = value;
        else if (name == "filter-type") { filterType = value; haveFilter = true; }
        else if (name == "practical-range") o.practicalRange = parseNum<float>(name, value);
        else if (name == "error-mean") o.errorMean = parseNum<float>(name, value);
        else if (name == "error-standard-deviation") o.errorStd = parseNum<float>(name, value);
        else if (name == "iterations") o.iterations = parseNum<int>(name, value);
        else if (name == "gpu") o.useGPU = hasValue ? parseBool(value) : true;
        else if (name == "cpu") o.useCPU = hasValue ? parseBool(value) : true;
        else if (name == "staged") o.staged = hasValue ? parseBool(value) : true;
        else if (name == "timing") o.timing = hasValue ? parseBool(value) : true;
        else if (name == "output-file") { outputFile = value; haveOutputFile = true; }
        else if (name == "output-option") { outputOption = value; haveOutputOption = true; }
        else if (name == "seed") { o.seed = parseNum<uint32_t>(name, value); o.seedGiven = true; }
        else if (name == "checkpoint") o.checkpoint = value;
        else if (name == "checkpoint-every") o.checkpointEvery = parseNum<int>(name, value);
        else if (name == "resume") o.resume = value;
        else if (name == "fast-convolution") o.fastConvolution = hasValue ? parseBool(value) : true;
        else if (name == "surrogate") o.surrogate = value;
        else if (name == "surrogate-check") o.surrogateCheck = parseNum<int>(name, value);
        else if (name == "surrogate-l1") o.surrogateL1 = parseNum<double>(name, value);
        else if (name == "surrogate-rms-ratio") o.surrogateRmsRatio = parseNum<double>(name, value);
        else if (name == "surrogate-window") o.surrogateWindow = parseNum<int>(name, value);
        else if (name == "surrogate-refine") o.surrogateRefine = parseNum<int>(name, value);
        else if (name == "batch") o.batchFile = value;
        else if (name == "batch-mode") o.batchMode = value;
        else if (name == "batch-tune") o.batchTune = hasValue ? parseBool(value) : true;
        else if (name == "synth") {
            int w = 0, h = 0;
            unsigned long long s = 0;
            if (std::sscanf(value.c_str(), "%d,%d,%llu", &w, &h, &s) != 3 || w <= 0 || h <= 0)
                fail("Bad value for --synth (W,H,SEED): " + value);
            o.synthW = w;
            o.synthH = h;
            o.synthSeed = s;
        } else {
            fail("unrecognised option '--" + name + "'");
        }
    }
    // main.cpp:383-403: the extension picks the format, the stem the name
    if (haveOutputFile) {
        std::string ext;
        const std::string name = fileStem(outputFile, &ext);
        if (ext == ".bin") o.jpeg = false;
        else if (ext == ".jpg" || ext.empty()) o.jpeg = true;
        else fail("Unknown output extension provided: " + ext + ". Exiting...");
        if (!name.empty()) o.outputName = name;
        o.outputOption = Options::OUT_END;
    }
    if (haveOutputOption) {
        if (outputOption == "end") o.outputOption = Options::OUT_END;
        else if (outputOption == "frames") o.outputOption = Options::OUT_FRAMES;
        else fail("Unknown output option provided: " + outputOption + ". Exiting...");
    }
    if (haveFilter) {
        if (filterType == "Exponential") o.filterModel = EXPONENTIAL;
        else if (filterType == "Gaussian") o.filterModel = GAUSSIAN;
        else fail("Unknown filter type provided: " + filterType + ". Exiting...");
    }
    if (o.useCPU) o.useGPU = false;
    if (o.iterations < 0) fail("The number of iterations must not be negative");
    if (o.checkpointEvery < 0) fail("--checkpoint-every must not be negative");
    if (o.surrogateCheck < 2 || o.surrogateCheck > 1000) fail("--surrogate-check must be in 2..1000");
    if (o.surrogateL1 < 0 || o.surrogateRmsRatio < 0) fail("--surrogate-l1 and --surrogate-rms-ratio must not be negative");
    if (o.surrogateWindow < 0) fail("--surrogate-window must not be negative");
    if (o.surrogateRefine < 1) fail("--surrogate-refine must be positive");
    if (!o.surrogate.empty() && o.staged) fail("--surrogate runs on the fused path, not with --staged");
    if (o.batchMode != "graph" && o.batchMode != "stream" && o.batchMode != "per-tile")
        fail("Unknown batch mode: " + o.batchMode + " (graph, stream or per-tile)");
    if (!o.batchFile.empty() && (o.staged || !o.surrogate.empty() || !o.resume.empty() ||
                                 !o.checkpoint.empty() || o.outputOption == Options::OUT_FRAMES))
        fail("--batch runs the fused batched path only (no --staged, --surrogate, --resume, "
             "--checkpoint or frames output)");
    return o;
}

/* ---- resume point: counter raster + next realization ------------------- */
struct ResumePoint {
    std::map<std::string, std::string> kv;
};

std::string checkpointKey(const Options& o, int W, int H) {
    std::ostringstream s;
    s << "width " << W << "\nheight " << H << "\nseed " << o.seed << "\nfilter_model "
      << (int)o.filterModel << "\npractical_range " << std::setprecision(9) << o.practicalRange
      << "\nerror_mean " << o.errorMean << "\nerror_standard_deviation " << o.errorStd
      << "\nbevel_radius " << o.bevelRadius << "\n";
    return s.str();
}

void writeCheckpoint(const Options& o, int W, int H, int next, CatchmentBorder_t* mc) {
    const float* data = mc->getDataConst();
    const std::string tmpBin = o.checkpoint + ".bin.tmp", tmpTxt = o.checkpoint + ".tmp";
    writeRaster(tmpBin, data, (size_t)W * H);
    {
        std::ofstream f(tmpTxt.c_str(), std::ios::trunc);
        f << "cageo-resume-point 1\n" << checkpointKey(o, W, H) << "next_realization " << next << "\n";
        if (!f) fail("Could not write the resume point " + tmpTxt);
    }
    if (std::rename(tmpBin.c_str(), (o.checkpoint + ".bin").c_str()) != 0 ||
        std::rename(tmpTxt.c_str(), o.checkpoint.c_str()) != 0)
        fail("Could not move the resume point into place: " + o.checkpoint);
}

int readCheckpoint(Options* o, int W, int H, CatchmentBorder_t* mc) {
    std::ifstream f(o->resume.c_str());
    if (!f.is_open()) fail("Unable to open the resume point \"" + o->resume + "\"");
    std::string magic, line;
    std::getline(f, magic);
    if (magic != "cageo-resume-point 1") fail("Not a resume point: " + o->resume);
    std::map<std::string, std::string> kv;
    while (std::getline(f, line)) {
        std::istringstream is(line);
        std::string k, v;
        if (is >> k >> v) kv[k] = v;
    }
    if (!kv.count("seed") || !kv.count("next_realization")) fail("Incomplete resume point");
    const uint32_t seed = (uint32_t)std::stoul(kv["seed"]);
    if (o->seedGiven && seed != o->seed) fail("--seed differs from the resume point's seed");
    o->seed = seed;
    o->seedGiven = true;
    std::istringstream want(checkpointKey(*o, W, H));
    std::string k, v;
    while (want >> k >> v)
        if (kv[k] != v) fail("The resume point was made with " + k + " " + kv[k] + ", not " + v);
    readRaster(o->resume + ".bin", mc->getData(), (size_t)W * H);
    return std::stoi(kv["next_realization"]);
}

/* ---- output: main.cpp:250-281 ---------------------------------------- */
void writeOutput(const Options& o, int iter, CatchmentBorder_t* mc,
                 ToByteArrayAlgorithm<float>* toBytes, Timer* byteTimer, Timer* jpegTimer) {
    std::ostringstream name;
    name << o.outputName;
    if (o.outputOption == Options::OUT_FRAMES) name << "_" << std::setfill('0') << std::setw(4) << iter;
    if (o.jpeg) {
        CellGrid<uint8_t> bytes(*mc, AS_TEMPLATE);
        byteTimer->startTimer("Converting to byte array");
        toBytes->execute(&bytes, mc);
        byteTimer->endTimer();
        jpegTimer->startTimer("JPEG Encode time");
        writeJpegGray(name.str() + ".jpg", bytes.getDataConst(), mc->width, mc->height);
        jpegTimer->endTimer();
    } else {
        writeRaster(name.str() + ".bin", mc->getDataConst(), (size_t)mc->size);
    }
}

// Surrogate prediction with its Monte Carlo check (cageo_surrogate_*,
// cageo_check_probability_map).  With one window: on acceptance `mc` holds
// probability x iterations (the expected border count of the run) and true is
// returned, otherwise the caller runs the Monte Carlo.  With windows, failing
// windows take the frequency of a --surrogate-refine Monte Carlo and true is
// returned.
bool runSurrogate(const Options& o, cageo_context* ctx, int W, int H, RasterDEM_t& dem,
                  StreamData_t& streams, CatchmentBorder_t* mc) {
    const auto t0 = std::chrono::steady_clock::now();
    cageo_surrogate* s = nullptr;
    check(cageo_surrogate_create(&s, o.surrogate.c_str(), nullptr), "surrogate weights");
    float *dProb = nullptr, *dScratch = nullptr;
    hip_check(hipMalloc(&dProb, (size_t)W * H * sizeof(float)), "hipMalloc");
    hip_check(hipMalloc(&dScratch, (size_t)W * H * sizeof(float)), "hipMalloc");
    check(cageo_surrogate_predict(s, dem.getDeviceDataConst(), streams.getDeviceDataConst(), W, H,
                                  (float)G::noDataValueDEM, (float)G::noDataValueStream, o.errorStd,
                                  o.practicalRange, dProb),
          "surrogate prediction");
    const double biasTol = 0.5 * o.surrogateL1, rmsTol = 1.2 * o.surrogateRmsRatio * o.surrogateL1;
    int nwx = 1, nwy = 1;
    check(cageo_check_windows(ctx, o.surrogateWindow, &nwx, &nwy), "check windows");
    std::vector<int> accepted((size_t)nwx * nwy);
    std::vector<double> bias(accepted.size()), rms(accepted.size());
    check(cageo_check_probability_map(ctx, dProb, o.surrogateCheck, (uint64_t)o.seed ^ 0x5eedull,
                                      o.surrogateWindow, biasTol, rmsTol, dScratch, accepted.data(),
                                      bias.data(), rms.data()),
          "surrogate check");
    int failed = 0;
    for (int a : accepted) failed += a == 0;
    std::cout << "Surrogate: " << o.surrogateCheck << "-realization check of " << accepted.size()
              << " window(s): ";
    if (accepted.size() == 1)
        std::cout << "mean error " << std::setprecision(5) << bias[0] << " (|.| <= " << biasTol
                  << "), RMS error " << rms[0] << " (<= " << rmsTol << ") -> "
                  << (failed ? "rejected" : "accepted");
    else
        std::cout << failed << " failed";
    bool done = failed == 0;
    if (failed && o.surrogateWindow > 0) {
        check(cageo_refine_probability_map(ctx, dProb, o.surrogateWindow, accepted.data(),
                                           o.surrogateRefine, (uint64_t)o.seed ^ 0x7e1eull,
                                           dScratch),
              "surrogate refinement");
        std::cout << ", refined by a " << o.surrogateRefine << "-realization Monte Carlo";
        done = true;
    } else if (failed) {
        std::cout << ", running the Monte Carlo";
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    std::cout << " (" << std::setprecision(1) << std::fixed << ms << " ms)" << std::endl;
    std::cout.unsetf(std::ios::fixed);
    if (done) {
        std::vector<float> p((size_t)W * H);
        hip_check(hipMemcpy(p.data(), dProb, p.size() * sizeof(float), hipMemcpyDeviceToHost),
                  "hipMemcpy");
        float* out = mc->getData();
        for (size_t i = 0; i < p.size(); ++i) out[i] = p[i] * (float)o.iterations;
    }
    (void)hipFree(dProb);
    (void)hipFree(dScratch);
    cageo_surrogate_destroy(s);
    return done;
}

void synthesize(const Options& o) {
    const int W = o.synthW, H = o.synthH;
    std::vector<float> dem((size_t)W * H), st((size_t)W * H);
    check(cageo_synth_dem(W, H, o.synthSeed, 10.0f, o.accThreshold, -9999.0f, dem.data(), st.data()),
          "synthetic DEM");
    DemInfo info;
    info.fileCols = W;
    info.fileRows = H;
    info.cellSize = 10.0f;
    info.noDataValue = -9999.0f;
    const std::string d = o.demPath + "/" + o.demName, s = o.demPath + "/" + o.streamsName;
    writeHeader(d + ".hdr", info);
    writeRaster(d + ".bin", dem.data(), dem.size());
    writeHeader(s + ".hdr", info);
    writeRaster(s + ".bin", st.data(), st.size());
    std::cout << "Synthetic inputs: " << d << ".{hdr,bin}, " << s << ".{hdr,bin} (seed "
              << o.synthSeed << ")" << std::endl;
}

/* ---- batched form: many tiles x realizations in one run -------------- */
struct BatchTileSpec {
    std::string dem, stream;
    float errorStd, practicalRange;
    FilterModel model;
};

std::vector<BatchTileSpec> readBatchFile(const Options& o) {
    std::ifstream f(o.batchFile.c_str());
    if (!f.is_open()) fail("Unable to open the batch file \"" + o.batchFile + "\"");
    std::vector<BatchTileSpec> tiles;
    std::string line;
    int lineNo = 0;
    while (std::getline(f, line)) {
        ++lineNo;
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream is(line);
        BatchTileSpec t{"", "", o.errorStd, o.practicalRange, o.filterModel};
        if (!(is >> t.dem)) continue;  // blank line
        std::string filter;
        if (!(is >> t.stream)) fail("Batch file line " + std::to_string(lineNo) + ": no stream name");
        if (is >> t.errorStd) {
            if (!(is >> t.practicalRange))
                fail("Batch file line " + std::to_string(lineNo) + ": error std without practical range");
            if (is >> filter) {
                if (filter == "Exponential") t.model = EXPONENTIAL;
                else if (filter == "Gaussian") t.model = GAUSSIAN;
                else fail("Batch file line " + std::to_string(lineNo) + ": unknown filter type " + filter);
            }
        }
        tiles.push_back(t);
    }
    if (tiles.empty()) fail("The batch file lists no tiles");
    return tiles;
}

int runBatch(Options& o) {
    const std::vector<BatchTileSpec> spec = readBatchFile(o);
    const int T = (int)spec.size();
    if (!o.seedGiven) o.seed = (uint32_t)std::time(nullptr);
    std::cout << "Batch: " << T << " tiles, " << o.iterations << " realizations each, mode "
              << o.batchMode << ", seed " << o.seed << std::endl;
    std::vector<int> widths(T), heights(T);
    std::vector<DemInfo> info(T);
    std::vector<float*> dDem(T, nullptr), dStream(T, nullptr), dMc(T, nullptr);
    for (int t = 0; t < T; ++t) {
        DemInfo si;
        loadHeaderData(o.demPath + "/" + spec[t].dem + ".hdr", &info[t]);
        loadHeaderData(o.demPath + "/" + spec[t].stream + ".hdr", &si);
        if (info[t].fileCols != si.fileCols || info[t].fileRows != si.fileRows)
            fail("The sizes of the DEM and stream data of tile " + spec[t].dem + " don't match");
        if (info[t].cellSize != si.cellSize)
            fail("The cell sizes of the DEM and stream data of tile " + spec[t].dem + " don't match");
        widths[t] = info[t].fileCols;
        heights[t] = info[t].fileRows;
        const size_t n = (size_t)widths[t] * heights[t];
        std::vector<float> dem(n), st(n);
        readRaster(o.demPath + "/" + spec[t].dem + ".bin", dem.data(), n);
        readRaster(o.demPath + "/" + spec[t].stream + ".bin", st.data(), n);
        hip_check(hipMalloc(&dDem[t], n * sizeof(float)), "hipMalloc");
        hip_check(hipMalloc(&dStream[t], n * sizeof(float)), "hipMalloc");
        hip_check(hipMalloc(&dMc[t], n * sizeof(float)), "hipMalloc");
        hip_check(hipMemcpy(dDem[t], dem.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemcpy(dStream[t], st.data(), n * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
        hip_check(hipMemset(dMc[t], 0, n * sizeof(float)), "hipMemset");
    }
    cageo_batch* b = nullptr;
    check(cageo_batch_create(&b, T, widths.data(), heights.data(), nullptr), "batch");
    for (int t = 0; t < T; ++t) {
        DemInfo si;
        loadHeaderData(o.demPath + "/" + spec[t].stream + ".hdr", &si);
        check(cageo_batch_setup_tile(b, t, dDem[t], dStream[t], info[t].noDataValue, si.noDataValue,
                                     o.bevelRadius, spec[t].model, spec[t].practicalRange,
                                     info[t].cellSize, o.errorMean, spec[t].errorStd,
                                     (uint64_t)o.seed + (uint64_t)t),
              ("batch tile " + spec[t].dem).c_str());
    }
    if (o.batchTune) {
        std::vector<float*> scratch(T);
        for (int t = 0; t < T; ++t) {
            hip_check(hipMalloc(&scratch[t], (size_t)widths[t] * heights[t] * sizeof(float)), "hipMalloc");
            hip_check(hipMemset(scratch[t], 0, (size_t)widths[t] * heights[t] * sizeof(float)), "hipMemset");
        }
        check(cageo_batch_tune(b, scratch.data(), 1, nullptr), "batch tune");
        for (float* p : scratch) (void)hipFree(p);
    }
    const int mode = o.batchMode == "graph" ? CAGEO_BATCH_GRAPH
                     : o.batchMode == "stream" ? CAGEO_BATCH_STREAM : CAGEO_BATCH_PER_TILE;
    const auto t0 = std::chrono::steady_clock::now();
    check(cageo_batch_run(b, dMc.data(), o.iterations, 0u, mode), "batch run");
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    double cells = 0.0;
    for (int t = 0; t < T; ++t) cells += (double)widths[t] * heights[t];
    std::cout << "Realizations: " << T << " tiles x " << o.iterations << " in " << std::fixed
              << std::setprecision(1) << ms << " ms (" << std::setprecision(2)
              << (ms > 0 ? 1000.0 * T * o.iterations / ms : 0.0) << " tile realizations/s, "
              << (ms > 0 ? 1e-3 * cells * o.iterations / ms : 0.0) << " Mcells/s)" << std::endl;
    std::cout.unsetf(std::ios::fixed);
    for (int t = 0; t < T; ++t) {
        const size_t n = (size_t)widths[t] * heights[t];
        const std::string name = o.outputName + "_" + spec[t].dem;
        if (o.jpeg) {
            cageo_context* ctx = nullptr;
            check(cageo_batch_context(b, t, &ctx), "batch context");
            uint8_t* dBytes = nullptr;
            hip_check(hipMalloc(&dBytes, n), "hipMalloc");
            check(cageo_to_byte_array(ctx, dBytes, dMc[t]), "byte array");
            std::vector<uint8_t> bytes(n);
            hip_check(hipMemcpy(bytes.data(), dBytes, n, hipMemcpyDeviceToHost), "hipMemcpy");
            (void)hipFree(dBytes);
            writeJpegGray(name + ".jpg", bytes.data(), widths[t], heights[t]);
        } else {
            std::vector<float> mc(n);
            hip_check(hipMemcpy(mc.data(), dMc[t], n * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
            writeRaster(name + ".bin", mc.data(), n);
        }
    }
    cageo_batch_destroy(b);
    for (int t = 0; t < T; ++t) {
        (void)hipFree(dDem[t]);
        (void)hipFree(dStream[t]);
        (void)hipFree(dMc[t]);
    }
    return 0;
}

int run(int argc, char** argv) {
    Options o = parseArgs(argc, argv);
    if (!o.useGPU) {
        std::cout << "Using OpenMP version" << std::endl;
        fail("The CPU implementation is the reference's own and is not built into this program; "
             "run without --cpu to use the GPU_HIP stages.");
    }
    AbstractFactory::defaultImpl = AbstractFactory::GPU_HIP;

    int devices = 0;
    if (hipGetDeviceCount(&devices) != hipSuccess || devices < 1) fail("No HIP device found");
    hip_check(hipSetDevice(0), "hipSetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, 0), "hipGetDeviceProperties");
    std::cout << "Acquiring device " << prop.name << " (" << prop.gcnArchName << ", "
              << prop.multiProcessorCount << " CUs, " << (prop.totalGlobalMem >> 30) << " GiB)"
              << std::endl;

    if (o.synthW > 0) synthesize(o);
    if (!o.batchFile.empty()) return runBatch(o);

    // main.cpp:87-108; the stream files are read from the DEM path
    DemInfo demInfo, streamInfo;
    loadHeaderData(o.demPath + "/" + o.demName + ".hdr", &demInfo);
    loadHeaderData(o.demPath + "/" + o.streamsName + ".hdr", &streamInfo);
    bool headerError = false;
    if (demInfo.fileCols != streamInfo.fileCols) {
        std::cout << "The sizes of the DEM and stream data don't match! Exiting..." << std::endl;
        headerError = true;
    }
    if (demInfo.cellSize != streamInfo.cellSize) {
        std::cout << "The cell size of the DEM and stream data don't match! Exiting..." << std::endl;
        headerError = true;
    }
    if (headerError) std::exit(1);
    const int W = demInfo.fileCols, H = demInfo.fileRows;
    if (W <= 0 || H <= 0) fail("Bad raster size in " + o.demName + ".hdr");
    if (o.bevelRadius > CAGEO_MAX_BEVEL_RADIUS)
        fail("Bevel radius is larger than MAX_BEVEL_RADIUS");
    G::noDataValueDEM = demInfo.noDataValue;
    G::noDataValueStream = streamInfo.noDataValue;
    const int filterRadius = (int)std::ceil((o.practicalRange / demInfo.cellSize) * 2.5f);
    std::cout << "DEM size: " << W << " x " << H << "\n";
    std::cout << "Kernel size: " << filterRadius * 2 + 1 << " x "
              << (o.filterModel == GAUSSIAN ? 1 : filterRadius * 2 + 1) << " ("
              << (o.filterModel == GAUSSIAN ? "Gaussian" : "Exponential") << ")\n"
              << std::endl;

    Timer mcTimer("Monte Carlo iteration"), globalTimer;
    Timer rfTimer("Random field"), convTimer("Convolution"), realTimer("Realistic DEM"),
        burnTimer("Stream burning"), pfTimer("Pit filling"), frTimer("Flow routing"),
        ftTimer("Flow tracing"), cbTimer("Catchment border extraction");
    Timer byteTimer("Converting to byte array"), jpegTimer("JPEG Encode time");
    globalTimer.startTimer("Total algorithm time");

    std::unique_ptr<RandomFieldAlgorithm_t> randomField(RandomFieldAlgorithmFactory::create<float>());
    randomField->setMean(o.errorMean);
    randomField->setStandardDeviation(o.errorStd);
    std::unique_ptr<ConvolutionAlgorithm_t> convolution(ConvolutionAlgorithmFactory::create<float>());
    convolution->setFilterRadius(filterRadius);
    convolution->setFilterModel(o.filterModel);
    convolution->createFilter(o.practicalRange, demInfo.cellSize);
    std::unique_ptr<RealisticDemAlgorithm_t> realistic(new RealisticDemAlgorithm_t());
    std::unique_ptr<StreamBurningAlgorithm_t> burning(StreamBurningAlgorithmFactory::create<DemData, DemData>());
    burning->setBevelRadius(o.bevelRadius);
    burning->setNoDataValueDEM((DemData)G::noDataValueDEM);
    burning->setNoDataValueStream((DemData)G::noDataValueStream);
    std::unique_ptr<PitFillingAlgorithm_t> pitFilling(PitFillingAlgorithmFactory::create<DemData>());
    pitFilling->setNoDataValueDem((DemData)G::noDataValueDEM);
    std::unique_ptr<FlowRoutingAlgorithm_t> routing(FlowRoutingAlgorithmFactory::create<cageo_host::char2, DemData>());
    routing->setNoDataValueDEM((DemData)G::noDataValueDEM);
    std::unique_ptr<FlowTracingAlgorithm_t> tracing(FlowTracingAlgorithmFactory::create<int, cageo_host::char2, DemData>());
    std::unique_ptr<CatchmentBorderAlgorithm_t> border(CatchmentBorderAlgorithmFactory::create<float, int>());
    std::unique_ptr<ToByteArrayAlgorithm<float>> toBytes(ToByteArrayAlgorithmFactory::create<float>());

    // inputs that stay for the whole run (main.cpp:161-178)
    RasterDEM_t dem(W, H);
    StreamData_t streams(W, H);
    readRaster(o.demPath + "/" + o.demName + ".bin", dem.getData(), (size_t)W * H);
    readRaster(o.demPath + "/" + o.streamsName + ".bin", streams.getData(), (size_t)W * H);
    CatchmentBorder_t monteCarloBorder(dem, AS_TEMPLATE);
    monteCarloBorder.fill(0.0f);

    int first = 0;
    if (!o.resume.empty()) {
        first = readCheckpoint(&o, W, H, &monteCarloBorder);
        std::cout << "Resuming at realization " << first << std::endl;
    }
    if (!o.seedGiven) o.seed = (uint32_t)std::time(nullptr);
    std::cout << "Seed: " << o.seed << std::endl;

    auto afterIteration = [&](int iter) {
        if (!o.checkpoint.empty() && o.checkpointEvery > 0 && (iter + 1) % o.checkpointEvery == 0 &&
            iter + 1 < o.iterations)
            writeCheckpoint(o, W, H, iter + 1, &monteCarloBorder);
        if ((o.outputOption == Options::OUT_END && iter + 1 == o.iterations) ||
            o.outputOption == Options::OUT_FRAMES)
            writeOutput(o, iter, &monteCarloBorder, toBytes.get(), &byteTimer, &jpegTimer);
    };

    if (o.staged) {
        for (int iter = first; iter < o.iterations; ++iter) {
            mcTimer.startTimer("Monte Carlo iteration");
            std::unique_ptr<RandomField_t> rand(new RandomField_t(dem, AS_TEMPLATE));
            rfTimer.startTimer("Random field");
            randomField->setRealization((uint32_t)iter);
            randomField->execute(rand.get(), (int)o.seed);
            rfTimer.endTimer();
            convTimer.startTimer("Convolution");
            convolution->execute(rand.get());
            convTimer.endTimer();
            RasterDEM_t* rDem = rand.get();  // the error field becomes the DEM in place
            realTimer.startTimer("Realistic DEM");
            realistic->execute(&dem, rand.get(), rDem);
            realTimer.endTimer();
            burnTimer.startTimer("Stream burning");
            burning->execute(rDem, &streams);
            burnTimer.endTimer();
            std::unique_ptr<RasterDEM_t> filled(new RasterDEM_t(*rDem, AS_TEMPLATE));
            pfTimer.startTimer("Pit filling");
            pitFilling->execute(filled.get(), rDem);
            pfTimer.endTimer();
            rand.reset();
            std::unique_ptr<FlowDir_t> flowDir(new FlowDir_t(dem, AS_TEMPLATE));
            frTimer.startTimer("Flow routing");
            routing->execute(flowDir.get(), filled.get());
            frTimer.endTimer();
            filled.reset();
            std::unique_ptr<FlowTraceData_t> trace(new FlowTraceData_t(*flowDir, AS_TEMPLATE));
            ftTimer.startTimer("Flow tracing");
            tracing->execute(trace.get(), flowDir.get(), &streams);
            ftTimer.endTimer();
            if (tracing->longPaths > 0)
                std::cout << tracing->longPaths << " cells with flow paths longer than 100001 steps"
                          << std::endl;
            flowDir.reset();
            cbTimer.startTimer("Catchment border extraction");
            border->execute(&monteCarloBorder, trace.get());
            cbTimer.endTimer();
            mcTimer.endTimer();
            if (o.timing) {
                for (const Timer* t : {&rfTimer, &convTimer, &realTimer, &burnTimer, &pfTimer,
                                       &frTimer, &ftTimer, &cbTimer, &mcTimer})
                    std::cerr << "[TIMING] " << t->getTimingTarget() << ": " << t->getLastTime()
                              << " ms" << std::endl;
            }
            afterIteration(iter);
        }
    } else {
        // The fused device loop: same stages, same buffers semantics, one
        // call per chunk of realizations (chunks end where output or a
        // resume point is due).
        HipContexts::Entry& e = HipContexts::get(W, H);
        check(cageo_set_filter(e.ctx, convolution->filterModel, convolution->filterRadius,
                               convolution->filter.data()),
              "set filter");
        e.model = convolution->filterModel;
        e.radius = convolution->filterRadius;
        e.taps = convolution->filter;
        check(cageo_mc_setup(e.ctx, dem.getDeviceDataConst(), streams.getDeviceDataConst(),
                             (float)G::noDataValueDEM, (float)G::noDataValueStream, o.bevelRadius,
                             o.errorMean, o.errorStd),
              "Monte Carlo set-up");
        if (o.fastConvolution)
            check(cageo_context_set_conv_mode(e.ctx, CAGEO_CONV_FAST), "fast convolution");
        check(cageo_stage_times(e.ctx, nullptr, o.timing ? 1 : 0), "stage timing");
        if (!o.surrogate.empty() && first < o.iterations && runSurrogate(o, e.ctx, W, H, dem, streams,
                                                                          &monteCarloBorder)) {
            writeOutput(o, o.iterations - 1, &monteCarloBorder, toBytes.get(), &byteTimer, &jpegTimer);
            first = o.iterations;  // no Monte Carlo left to run
        }
        Timer* stageTimers[7] = {&rfTimer, &convTimer, &realTimer, &pfTimer, &frTimer, &ftTimer, &cbTimer};
        realTimer.setTimingTarget("Realistic DEM + stream burning");
        int iter = first;
        while (iter < o.iterations) {
            int chunk = o.iterations - iter;
            if (o.outputOption == Options::OUT_FRAMES) chunk = 1;
            if (!o.checkpoint.empty() && o.checkpointEvery > 0)
                chunk = std::min(chunk, o.checkpointEvery - iter % o.checkpointEvery);
            const auto t0 = std::chrono::steady_clock::now();
            check(cageo_run_realizations(e.ctx, monteCarloBorder.getDeviceData(), chunk,
                                         (uint64_t)o.seed, (uint32_t)iter),
                  "Monte Carlo realizations");
            int64_t longPaths = 0;
            check(cageo_context_long_paths(e.ctx, &longPaths), "long paths");
            if (longPaths > 0)
                std::cout << longPaths << " cells with flow paths longer than 100001 steps in "
                          << chunk << " realizations" << std::endl;
            // one interval covers `chunk` realizations
            mcTimer.add(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                            .count(),
                        chunk);
            if (o.timing) {
                float ms[7];
                check(cageo_stage_times(e.ctx, ms, 1), "stage timing");
                for (int s = 0; s < 7; ++s) stageTimers[s]->add(ms[s], chunk);
            }
            iter += chunk;
            afterIteration(iter - 1);
        }
    }
    if (!o.checkpoint.empty()) writeCheckpoint(o, W, H, o.iterations, &monteCarloBorder);
    globalTimer.endTimer();

    const int done = o.iterations - first;
    std::cout << "Realizations: " << done << " in " << std::fixed << std::setprecision(1)
              << mcTimer.getTotalTime() << " ms ("
              << std::setprecision(2) << (mcTimer.getTotalTime() > 0 ? 1000.0 * done / mcTimer.getTotalTime() : 0.0)
              << " realizations/s, "
              << (mcTimer.getTotalTime() > 0 ? 1e-3 * W * (double)H * done / mcTimer.getTotalTime() : 0.0)
              << " Mcells/s)" << std::endl;
    if (o.timing) {
        std::cerr << "\n\n***********\n\n";
        std::cerr << "[TIMING] " << globalTimer.getTimingTarget() << ": " << globalTimer.getTotalTime()
                  << " ms (total)" << std::endl;
        std::vector<const Timer*> rows = {&mcTimer, &rfTimer, &convTimer, &realTimer};
        if (o.staged) rows.push_back(&burnTimer);
        for (const Timer* t : {&pfTimer, &frTimer, &ftTimer, &cbTimer}) rows.push_back(t);
        for (const Timer* t : rows)
            std::cerr << "[TIMING] " << t->getTimingTarget() << ": " << t->getAverageTime()
                      << " ms (average)" << std::endl;
        TimerTree tree(&globalTimer);
        for (size_t i = 1; i < rows.size(); ++i) tree.setChild(rows[i]);
        tree.summary(std::cout);
    }
    HipContexts::releaseAll();
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& ex) {
        std::cout << "drainage: " << ex.what() << std::endl;
        return 1;
    }
}
