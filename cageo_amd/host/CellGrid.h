/*
 * Device-resident raster grids for the drainage pipeline on MI355X.
 *
 * Mirrors the reference's CellGrid<T> (framework/CellGrid.h) over
 * ManagedData<T> (framework/ManagedData.h:502-773, ManagedDataCUDA.h:138-213):
 * a W x H row-major grid whose data lives on the host, the device, or both,
 * migrated on demand by toHost()/toDevice().  The reference also spills to
 * disk when host memory runs short; a whole 16384^2 float pipeline needs a few
 * GiB against 288 GB of HBM per GPU, so here the device copy is the primary
 * one and there is no disk tier.
 *
 * A grid made "AS_TEMPLATE" from another takes its shape only (the
 * reference's CellGrid(const CellGrid&, CopyType) with AS_TEMPLATE).
 */
#ifndef CAGEO_HOST_CELLGRID_H_
#define CAGEO_HOST_CELLGRID_H_

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

namespace cageo_host {

/* char2 {x, y} flow direction of FlowDirClass.h:44 as a plain POD. */
struct char2 {
    int8_t x, y;
};
static_assert(sizeof(char2) == 2, "char2 must be two bytes");

enum CopyType { AS_TEMPLATE, DEEP_COPY };

/* Where the newest copy of a grid lives. */
enum Residence { NONE, HOST, DEVICE, BOTH };

inline void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

template <typename T>
class ManagedData {
public:
    explicit ManagedData(size_t count) : count_(count) {}
    ~ManagedData() {
        if (dev_) (void)hipFree(dev_);
    }
    ManagedData(const ManagedData&) = delete;
    ManagedData& operator=(const ManagedData&) = delete;

    size_t count() const { return count_; }
    Residence residence() const { return where_; }

    /* Host pointer with the newest data (copied down if the device is newer). */
    T* toHost() {
        if (host_.size() != count_) host_.assign(count_, T());
        if (where_ == DEVICE) {
            hip_check(hipMemcpy(host_.data(), dev_, bytes(), hipMemcpyDeviceToHost), "toHost");
            where_ = BOTH;
        } else if (where_ == NONE) {
            where_ = HOST;
        }
        return host_.data();
    }
    /* Device pointer with the newest data (copied up if the host is newer). */
    T* toDevice() {
        if (!dev_) hip_check(hipMalloc(reinterpret_cast<void**>(&dev_), bytes() ? bytes() : 1),
                             "hipMalloc");
        if (where_ == HOST) {
            hip_check(hipMemcpy(dev_, host_.data(), bytes(), hipMemcpyHostToDevice), "toDevice");
            where_ = BOTH;
        } else if (where_ == NONE) {
            where_ = DEVICE;
        }
        return dev_;
    }
    /* Record that the host copy (resp. device copy) was written. */
    void hostModified() { where_ = HOST; }
    void deviceModified() { where_ = DEVICE; }
    /* Give the host copy back (the device copy stays authoritative). */
    void releaseHost() {
        if (where_ == BOTH) where_ = DEVICE;
        if (where_ == DEVICE) std::vector<T>().swap(host_);
    }
    size_t bytes() const { return count_ * sizeof(T); }

private:
    size_t count_;
    std::vector<T> host_;
    T* dev_ = nullptr;
    Residence where_ = NONE;
};

template <typename T>
class CellGrid {
public:
    int width = 0, height = 0, size = 0;

    CellGrid(int w, int h) : width(w), height(h), size(w * h), data_(new ManagedData<T>((size_t)w * h)) {
        if (w <= 0 || h <= 0) throw std::invalid_argument("CellGrid: empty grid");
    }
    template <typename U>
    CellGrid(const CellGrid<U>& other, CopyType type)
        : CellGrid(other.width, other.height) {
        if (type != AS_TEMPLATE) throw std::invalid_argument("CellGrid: only AS_TEMPLATE across types");
    }
    CellGrid(const CellGrid& other, CopyType type) : CellGrid(other.width, other.height) {
        if (type == DEEP_COPY) {
            CellGrid& src = const_cast<CellGrid&>(other);
            hip_check(hipMemcpy(getDeviceData(), src.getDeviceDataConst(), data_->bytes(),
                                hipMemcpyDeviceToDevice),
                      "deep copy");
            data_->deviceModified();
        }
    }
    virtual ~CellGrid() { delete data_; }
    CellGrid(const CellGrid&) = delete;
    CellGrid& operator=(const CellGrid&) = delete;

    void toHost() { (void)data_->toHost(); }
    void toDevice() { (void)data_->toDevice(); }
    /* Host data, newest copy; the caller is about to write it. */
    T* getData() {
        T* p = data_->toHost();
        data_->hostModified();
        return p;
    }
    /* Host data for reading only. */
    const T* getDataConst() const { return data_->toHost(); }
    /* Device data, newest copy; the caller (a stage) is about to write it. */
    T* getDeviceData() {
        T* p = data_->toDevice();
        data_->deviceModified();
        return p;
    }
    /* Device data for reading only. */
    const T* getDeviceDataConst() { return data_->toDevice(); }
    Residence residence() const { return data_->residence(); }
    void releaseHost() { data_->releaseHost(); }

    /* Fill every cell with `v` (host side, then the device copy is stale). */
    void fill(T v) {
        T* p = getData();
        for (int i = 0; i < size; ++i) p[i] = v;
    }

private:
    ManagedData<T>* data_;
};

/* Data classes of the reference by role.  They are one type, so that the
 * random field can be used as the DEM with error in place as main.cpp:197
 * does with a cast. */
template <typename T> using RasterDEM = CellGrid<T>;
template <typename T> using StreamData = CellGrid<T>;
template <typename T> using RandomFieldClass = CellGrid<T>;
template <typename T> using FlowDirClass = CellGrid<T>;
template <typename T> using FlowTraceDataClass = CellGrid<T>;
template <typename T> using CatchmentBorderClass = CellGrid<T>;

typedef float DemData;
typedef RasterDEM<DemData> RasterDEM_t;
typedef StreamData<DemData> StreamData_t;
typedef RandomFieldClass<float> RandomField_t;
typedef FlowDirClass<char2> FlowDir_t;
typedef FlowTraceDataClass<int> FlowTraceData_t;
typedef CatchmentBorderClass<float> CatchmentBorder_t;

}  // namespace cageo_host

#endif  // CAGEO_HOST_CELLGRID_H_
