/*
 * The reference's stage plugin API with an MI355X implementation behind it.
 *
 * Each stage keeps the reference's abstract class (public execute(),
 * protected pure virtual perform*()) and setters, and adds a *_GPU_HIP
 * subclass whose perform*() is a thin call into libcageo_hip's C ABI
 * (include/cageo_hip.h).  The factories switch on AbstractFactory::defaultImpl
 * as the reference's do (framework/AbstractFactory.h:31-41); GPU and GPU_HIP
 * give the HIP implementation.  The CPU and CUDA implementations are the
 * reference's own classes and are not part of this library: asking for them
 * raises an error instead of silently running something else.
 *
 * Errors: a non-zero status from the C ABI becomes a StageError carrying the
 * status; the drainage CLI prints it and exits non-zero (the reference prints
 * and exits on configuration errors, main.cpp:78-100).
 */
#ifndef CAGEO_HOST_STAGES_H_
#define CAGEO_HOST_STAGES_H_

#include <cmath>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "CellGrid.h"
#include "cageo_hip.h"

namespace cageo_host {

/* FilterModel, ConvolutionCommon.h:18-21. */
enum FilterModel { EXPONENTIAL = CAGEO_FILTER_EXPONENTIAL, GAUSSIAN = CAGEO_FILTER_GAUSSIAN };

/* Globals of GlobalParameters.cpp used by the stages (RealisticDemAlgorithm
 * reads G::noDataValueDEM, RealisticDemAlgorithm.h:59-60). */
namespace G {
extern double noDataValueDEM;
extern double noDataValueStream;
}  // namespace G

class StageError : public std::runtime_error {
public:
    StageError(const std::string& what, int status)
        : std::runtime_error(what + ": " + cageo_status_string(status)), status(status) {}
    int status;
};

inline void check(int status, const char* what) {
    if (status != CAGEO_OK) throw StageError(what, status);
}

/* One libcageo_hip context (device workspace) per grid shape, shared by the
 * stages of this process, plus the filter last bound to it. */
class HipContexts {
public:
    struct Entry {
        cageo_context* ctx = nullptr;
        int model = -1, radius = -1;
        std::vector<float> taps;
    };
    static Entry& get(int width, int height) {
        auto& m = map();
        auto it = m.find({width, height});
        if (it != m.end()) return *it->second;
        std::unique_ptr<Entry> e(new Entry());
        check(cageo_context_create(&e->ctx, width, height, nullptr), "cageo_context_create");
        Entry& ref = *e;
        m[{width, height}] = std::move(e);
        return ref;
    }
    static void releaseAll() {
        for (auto& kv : map()) (void)cageo_context_destroy(kv.second->ctx);
        map().clear();
    }

private:
    static std::map<std::pair<int, int>, std::unique_ptr<Entry>>& map() {
        static std::map<std::pair<int, int>, std::unique_ptr<Entry>> m;
        return m;
    }
};

template <typename T>
cageo_context* contextOf(CellGrid<T>* g) {
    return HipContexts::get(g->width, g->height).ctx;
}

class AbstractAlgorithm {
public:
    virtual ~AbstractAlgorithm() {}
};

class AbstractFactory {
public:
    /* AbstractFactory.h:31-36 plus GPU_HIP; GPU maps to GPU_HIP here. */
    enum Type { CPU, CPU_OpenMP, GPU, GPU_CUDA, GPU_HIP };
    static Type defaultImpl;
    static bool usesHip() { return defaultImpl == GPU || defaultImpl == GPU_HIP; }
    static void requireHip(const char* stage) {
        if (!usesHip())
            throw std::runtime_error(std::string(stage) +
                                     ": only the GPU_HIP implementation is built into this library");
    }
};

/* ---- random field: RandomFieldAlgorithm.h:40-108 ------------------------ */
template <typename T>
class RandomFieldAlgorithm : public AbstractAlgorithm {
public:
    float mean = 0.0f, standardDeviation = 1.0f;
    void execute(RandomFieldClass<T>* field, int seed) { performRandomField(field, seed); }
    void setMean(float m) { mean = m; }
    void setStandardDeviation(float s) { standardDeviation = s; }
    /* Philox counter word for the realization index (no reference
     * counterpart: the reference reseeds with time() + iteration). */
    void setRealization(uint32_t r) { realization = r; }

protected:
    uint32_t realization = 0;
    virtual void performRandomField(RandomFieldClass<T>* field, int seed) = 0;
};

template <typename T>
class RandomFieldAlgorithm_GPU_HIP : public RandomFieldAlgorithm<T> {
protected:
    void performRandomField(RandomFieldClass<T>* field, int seed) override {
        check(cageo_random_field(contextOf(field), field->getDeviceData(), (uint64_t)(uint32_t)seed,
                                 this->realization, 0u, this->mean, this->standardDeviation),
              "random field");
    }
};

/* ---- convolution: ConvolutionAlgorithm.h:40-170, .cpp:18-172 ----------- */
template <typename T>
class ConvolutionAlgorithm : public AbstractAlgorithm {
public:
    int filterRadius = 0;
    FilterModel filterModel = GAUSSIAN;
    std::vector<float> filter;

    void execute(RandomFieldClass<T>* field) { performConvolution(field); }
    void setFilterModel(FilterModel m) { filterModel = m; }
    void setFilterRadius(int r) { filterRadius = r; }
    /* createFilter(practicalRange, cellSize), ConvolutionAlgorithm.cpp:18-45. */
    void createFilter(float practicalRange, float cellSize) {
        const int L = 2 * filterRadius + 1;
        filter.assign(filterModel == GAUSSIAN ? (size_t)L : (size_t)L * L, 0.0f);
        check(cageo_create_filter(filterModel, practicalRange, cellSize, filterRadius,
                                  filter.data()),
              "createFilter");
    }

protected:
    virtual void performConvolution(RandomFieldClass<T>* field) = 0;
};

template <typename T>
class ConvolutionAlgorithm_GPU_HIP : public ConvolutionAlgorithm<T> {
protected:
    void performConvolution(RandomFieldClass<T>* field) override {
        if (this->filter.empty()) check(CAGEO_ERR_STATE, "convolution without createFilter");
        HipContexts::Entry& e = HipContexts::get(field->width, field->height);
        if (e.model != this->filterModel || e.radius != this->filterRadius || e.taps != this->filter) {
            check(cageo_set_filter(e.ctx, this->filterModel, this->filterRadius, this->filter.data()),
                  "set filter");
            e.model = this->filterModel;
            e.radius = this->filterRadius;
            e.taps = this->filter;
        }
        check(cageo_convolve(e.ctx, field->getDeviceData()), "convolution");
    }
};

/* ---- realistic DEM: RealisticDemAlgorithm.h:26-70 (concrete class) ------ */
template <typename T, typename U>
class RealisticDemAlgorithm : public AbstractAlgorithm {
public:
    /* demWithError may be the random field itself (main.cpp:200-204). */
    void execute(RasterDEM<T>* dem, RandomFieldClass<U>* field, RasterDEM<T>* demWithError) {
        AbstractFactory::requireHip("RealisticDemAlgorithm");
        float* e = field->getDeviceData();
        if ((void*)demWithError != (void*)field) {
            check(hipMemcpy(demWithError->getDeviceData(), e, (size_t)field->size * sizeof(float),
                            hipMemcpyDeviceToDevice) == hipSuccess
                      ? CAGEO_OK
                      : CAGEO_ERR_HIP,
                  "realistic DEM copy");
            e = demWithError->getDeviceData();
        }
        check(cageo_realistic_dem(contextOf(dem), dem->getDeviceDataConst(), e,
                                  (float)G::noDataValueDEM),
              "realistic DEM");
    }
};

/* ---- stream burning: StreamBurningAlgorithm.h:30-167 -------------------- */
template <typename T, typename U>
class StreamBurningAlgorithm : public AbstractAlgorithm {
public:
    int bevelRadius = 2;
    T noDataValueDEM = 0;
    U noDataValueStream = 0;
    void execute(RasterDEM<T>* demWithError, StreamData<U>* streams) {
        if (bevelRadius > CAGEO_MAX_BEVEL_RADIUS || bevelRadius < 0)
            check(CAGEO_ERR_ARG, "bevel radius above MAX_BEVEL_RADIUS");
        performStreamBurning(demWithError, demWithError, streams);
    }
    void setBevelRadius(int r) { bevelRadius = r; }
    void setNoDataValueDEM(T v) { noDataValueDEM = v; }
    void setNoDataValueStream(U v) { noDataValueStream = v; }

protected:
    virtual void performStreamBurning(RasterDEM<T>* out, RasterDEM<T>* dem,
                                      StreamData<U>* streams) = 0;
};

template <typename T, typename U>
class StreamBurningAlgorithm_GPU_HIP : public StreamBurningAlgorithm<T, U> {
protected:
    void performStreamBurning(RasterDEM<T>* out, RasterDEM<T>* /*dem: same grid*/,
                              StreamData<U>* streams) override {
        check(cageo_stream_burn(contextOf(out), out->getDeviceData(), streams->getDeviceDataConst(),
                                this->bevelRadius, this->noDataValueDEM, this->noDataValueStream),
              "stream burning");
    }
};

/* ---- pit filling: PitFillingAlgorithm.h:24-98 --------------------------- */
template <typename T>
class PitFillingAlgorithm : public AbstractAlgorithm {
public:
    int iter = 0;
    T noDataValueDEM = 0;
    /* dem is quantized to mm in place, as the reference does. */
    void execute(RasterDEM<T>* demPitsFilled, RasterDEM<T>* demWithStreams) {
        performPitFilling(demPitsFilled, demWithStreams);
    }
    void setNoDataValueDem(T v) { noDataValueDEM = v; }

protected:
    virtual void performPitFilling(RasterDEM<T>* spill, RasterDEM<T>* dem) = 0;
};

template <typename T>
class PitFillingAlgorithm_GPU_HIP : public PitFillingAlgorithm<T> {
protected:
    void performPitFilling(RasterDEM<T>* spill, RasterDEM<T>* dem) override {
        check(cageo_pit_fill(contextOf(dem), spill->getDeviceData(), dem->getDeviceData(),
                             this->noDataValueDEM),
              "pit filling");
    }
};

/* ---- flow routing: FlowRoutingAlgorithm.h:24-94 ------------------------- */
template <typename T, typename U>
class FlowRoutingAlgorithm : public AbstractAlgorithm {
public:
    U noDataValueDEM = 0;
    void setNoDataValueDEM(U v) { noDataValueDEM = v; }
    void execute(FlowDirClass<T>* flowDir, RasterDEM<U>* demPitsFilled) {
        performFlowRouting(flowDir, demPitsFilled);
        performFlowRoutingOnFlatSurfaces(flowDir, demPitsFilled);
    }

protected:
    virtual void performFlowRouting(FlowDirClass<T>* flowDir, RasterDEM<U>* dem) = 0;
    virtual void performFlowRoutingOnFlatSurfaces(FlowDirClass<T>* flowDir, RasterDEM<U>* dem) = 0;
};

template <typename T, typename U>
class FlowRoutingAlgorithm_GPU_HIP : public FlowRoutingAlgorithm<T, U> {
protected:
    /* D8 and the flat resolution run in one persistent launch. */
    void performFlowRouting(FlowDirClass<T>* flowDir, RasterDEM<U>* dem) override {
        static_assert(sizeof(T) == 2, "flow directions are char2");
        check(cageo_flow_routing(contextOf(dem), reinterpret_cast<int8_t*>(flowDir->getDeviceData()),
                                 dem->getDeviceDataConst(), this->noDataValueDEM),
              "flow routing");
    }
    void performFlowRoutingOnFlatSurfaces(FlowDirClass<T>*, RasterDEM<U>*) override {}
};

/* ---- flow tracing: FlowTracingAlgorithm.h:24-69 ------------------------- */
template <typename T, typename U, typename V>
class FlowTracingAlgorithm : public AbstractAlgorithm {
public:
    /* Cells whose path is longer than the reference's 100 001-step walk guard
     * in the last call (the reference prints a message for each). */
    long long longPaths = 0;
    void execute(FlowTraceDataClass<T>* trace, FlowDirClass<U>* flowDir, StreamData<V>* streams) {
        performFlowTracing(trace, flowDir, streams);
        performConnectFlowPath(trace);
    }

protected:
    virtual void performFlowTracing(FlowTraceDataClass<T>* trace, FlowDirClass<U>* flowDir,
                                    StreamData<V>* streams) = 0;
    virtual void performConnectFlowPath(FlowTraceDataClass<T>* trace) = 0;
};

template <typename T, typename U, typename V>
class FlowTracingAlgorithm_GPU_HIP : public FlowTracingAlgorithm<T, U, V> {
protected:
    /* Tracing and the connect step run in one persistent launch. */
    void performFlowTracing(FlowTraceDataClass<T>* trace, FlowDirClass<U>* flowDir,
                            StreamData<V>* streams) override {
        int64_t lp = 0;
        check(cageo_flow_tracing(contextOf(trace), reinterpret_cast<int32_t*>(trace->getDeviceData()),
                                 reinterpret_cast<const int8_t*>(flowDir->getDeviceDataConst()),
                                 streams->getDeviceDataConst(), &lp),
              "flow tracing");
        this->longPaths = lp;
    }
    void performConnectFlowPath(FlowTraceDataClass<T>*) override {}
};

/* ---- catchment border: CatchmentBorderAlgorithm.h:26-74 ----------------- */
template <typename T, typename U>
class CatchmentBorderAlgorithm : public AbstractAlgorithm {
public:
    /* monteCarloBorder += border, skipping cells equal to -1. */
    void execute(CatchmentBorderClass<T>* monteCarloBorder, FlowTraceDataClass<U>* trace) {
        performExtractDrainageBorder(monteCarloBorder, trace);
    }

protected:
    virtual void performExtractDrainageBorder(CatchmentBorderClass<T>* mc,
                                              FlowTraceDataClass<U>* trace) = 0;
};

template <typename T, typename U>
class CatchmentBorderAlgorithm_GPU_HIP : public CatchmentBorderAlgorithm<T, U> {
protected:
    void performExtractDrainageBorder(CatchmentBorderClass<T>* mc, FlowTraceDataClass<U>* trace) override {
        check(cageo_catchment_border(contextOf(mc), mc->getDeviceData(),
                                     reinterpret_cast<const int32_t*>(trace->getDeviceDataConst())),
              "catchment border");
    }
};

/* ---- visualisation: ToByteArrayAlgorithm.h ------------------------------ */
template <typename T>
class ToByteArrayAlgorithm : public AbstractAlgorithm {
public:
    void execute(CellGrid<uint8_t>* output, CellGrid<T>* input) { convertToByteArray(output, input); }

protected:
    virtual void convertToByteArray(CellGrid<uint8_t>* output, CellGrid<T>* input) = 0;
};

template <typename T>
class ToByteArrayAlgorithm_GPU_HIP : public ToByteArrayAlgorithm<T> {
protected:
    void convertToByteArray(CellGrid<uint8_t>* output, CellGrid<T>* input) override {
        check(cageo_to_byte_array(contextOf(input), output->getDeviceData(),
                                  input->getDeviceDataConst()),
              "to byte array");
    }
};

/* ---- factories: *AlgorithmFactory.h, create<...>() ---------------------- */
struct RandomFieldAlgorithmFactory {
    template <typename T> static RandomFieldAlgorithm<T>* create() {
        AbstractFactory::requireHip("RandomFieldAlgorithmFactory");
        return new RandomFieldAlgorithm_GPU_HIP<T>();
    }
};
struct ConvolutionAlgorithmFactory {
    template <typename T> static ConvolutionAlgorithm<T>* create() {
        AbstractFactory::requireHip("ConvolutionAlgorithmFactory");
        return new ConvolutionAlgorithm_GPU_HIP<T>();
    }
};
struct StreamBurningAlgorithmFactory {
    template <typename T, typename U> static StreamBurningAlgorithm<T, U>* create() {
        AbstractFactory::requireHip("StreamBurningAlgorithmFactory");
        return new StreamBurningAlgorithm_GPU_HIP<T, U>();
    }
};
struct PitFillingAlgorithmFactory {
    template <typename T> static PitFillingAlgorithm<T>* create() {
        AbstractFactory::requireHip("PitFillingAlgorithmFactory");
        return new PitFillingAlgorithm_GPU_HIP<T>();
    }
};
struct FlowRoutingAlgorithmFactory {
    template <typename T, typename U> static FlowRoutingAlgorithm<T, U>* create() {
        AbstractFactory::requireHip("FlowRoutingAlgorithmFactory");
        return new FlowRoutingAlgorithm_GPU_HIP<T, U>();
    }
};
struct FlowTracingAlgorithmFactory {
    template <typename T, typename U, typename V> static FlowTracingAlgorithm<T, U, V>* create() {
        AbstractFactory::requireHip("FlowTracingAlgorithmFactory");
        return new FlowTracingAlgorithm_GPU_HIP<T, U, V>();
    }
};
struct CatchmentBorderAlgorithmFactory {
    template <typename T, typename U> static CatchmentBorderAlgorithm<T, U>* create() {
        AbstractFactory::requireHip("CatchmentBorderAlgorithmFactory");
        return new CatchmentBorderAlgorithm_GPU_HIP<T, U>();
    }
};
struct ToByteArrayAlgorithmFactory {
    template <typename T> static ToByteArrayAlgorithm<T>* create() {
        AbstractFactory::requireHip("ToByteArrayAlgorithmFactory");
        return new ToByteArrayAlgorithm_GPU_HIP<T>();
    }
};

typedef RandomFieldAlgorithm<float> RandomFieldAlgorithm_t;
typedef ConvolutionAlgorithm<float> ConvolutionAlgorithm_t;
typedef RealisticDemAlgorithm<DemData, DemData> RealisticDemAlgorithm_t;
typedef StreamBurningAlgorithm<DemData, DemData> StreamBurningAlgorithm_t;
typedef PitFillingAlgorithm<DemData> PitFillingAlgorithm_t;
typedef FlowRoutingAlgorithm<char2, DemData> FlowRoutingAlgorithm_t;
typedef FlowTracingAlgorithm<int, char2, DemData> FlowTracingAlgorithm_t;
typedef CatchmentBorderAlgorithm<float, int> CatchmentBorderAlgorithm_t;

}  // namespace cageo_host

#endif  // CAGEO_HOST_STAGES_H_
