/*
 * diplib_amd.hpp -- C++ drop-in for DIPlib's Gaussian FIR path on MI355X.
 *
 * Keeps the reference signatures so existing callers compile unchanged:
 *   dip::Gauss                 (include/diplib/linear.h:452-472)
 *   dip::GaussFIR              (include/diplib/linear.h:300-318)
 *   dip::SeparableConvolution  (include/diplib/linear.h:114-130)
 *   dip::MakeHalfGaussian      (include/diplib/generation.h:638-653)
 * over the subset of dip::Image the path uses (include/diplib/library/image.h):
 * sizes, strides in samples (dim 0 fastest), tensor images (elements stored as planes;
 * filters treat each element as a scalar image, as the reference's AsScalarImage),
 * data type, a shared data segment, Forge / ReForge / Strip / Protect, and the
 * external-data constructor (image.h:458-491).  Pixel data live in DEVICE memory by
 * default (Forge allocates with hipMalloc); an image in HOST memory (MemoryLocation::
 * HOST: pinned by Forge, or the caller's own buffer) is copied to the device, filtered
 * there and its result copied back, so host callers need no other change.  Device
 * calls run asynchronously on the stream set with dipamd::SetStream (the null stream
 * by default), through the C ABI of include/dipamd.h; calls on host images return when
 * the result is in place.
 *
 * Output rules follow Framework::Separable (framework_separable.cpp:126): `out` is
 * reforged to the input's sizes and SuggestFlex(in) type unless it is protected,
 * in which case it keeps its type (written with clamp_cast) and must already have
 * the input's sizes ("Image is protected").  `out` may be `in` (in-place).
 * Errors are thrown as dip::ParameterError / dip::RunTimeError.
 *
 * Header-only C++14; link libdipamd.so and include ROCm's hip_runtime_api.h path.
 */
#ifndef DIPLIB_AMD_HPP
#define DIPLIB_AMD_HPP

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "dipamd.h"

namespace dipamd {
inline hipStream_t &CurrentStream() {
    thread_local hipStream_t s = nullptr;
    return s;
}
/* Stream for the dip:: calls of this thread. */
inline void SetStream(hipStream_t s) { CurrentStream() = s; }
}  // namespace dipamd

namespace dip {

using uint = std::size_t;
using sint = std::ptrdiff_t;
using dfloat = double;
using sfloat = float;
using String = std::string;
using StringArray = std::vector<String>;
using FloatArray = std::vector<dfloat>;
using UnsignedArray = std::vector<uint>;
using IntegerArray = std::vector<sint>;
using BooleanArray = std::vector<bool>;

/* error.h:53-123 */
class Error : public std::exception {
  public:
    explicit Error(std::string msg) : msg_(std::move(msg)) {}
    const char *what() const noexcept override { return msg_.c_str(); }

  private:
    std::string msg_;
};
class ParameterError : public Error {
  public:
    using Error::Error;
};
class RunTimeError : public Error {
  public:
    using Error::Error;
};

/* The sample types of datatype.h (without the 64-bit integers). */
class DataType {
  public:
    enum class DT : int32_t {
        UINT8 = DIPAMD_DT_UINT8, UINT16 = DIPAMD_DT_UINT16, UINT32 = DIPAMD_DT_UINT32, SINT8 = DIPAMD_DT_SINT8,
        SINT16 = DIPAMD_DT_SINT16, SINT32 = DIPAMD_DT_SINT32, SFLOAT = DIPAMD_DT_SFLOAT, DFLOAT = DIPAMD_DT_DFLOAT,
        BIN = DIPAMD_DT_BIN, SCOMPLEX = DIPAMD_DT_SCOMPLEX, DCOMPLEX = DIPAMD_DT_DCOMPLEX
    };
    constexpr DataType() : dt_(DT::SFLOAT) {}
    constexpr DataType(DT dt) : dt_(dt) {}  // NOLINT
    constexpr bool operator==(DataType o) const { return dt_ == o.dt_; }
    constexpr bool operator!=(DataType o) const { return dt_ != o.dt_; }
    constexpr int32_t Code() const { return static_cast<int32_t>(dt_); }
    std::size_t SizeOf() const {
        switch (dt_) {
            case DT::UINT8: case DT::SINT8: case DT::BIN: return 1;
            case DT::UINT16: case DT::SINT16: return 2;
            case DT::DFLOAT: case DT::SCOMPLEX: return 8;
            case DT::DCOMPLEX: return 16;
            default: return 4;
        }
    }
    bool IsComplex() const { return dt_ == DT::SCOMPLEX || dt_ == DT::DCOMPLEX; }
    /* datatype.cpp:109-124 */
    static DataType SuggestFlex(DataType t) {
        if (t.IsComplex()) return t;
        return (t.dt_ == DT::UINT32 || t.dt_ == DT::SINT32 || t.dt_ == DT::DFLOAT) ? DataType(DT::DFLOAT)
                                                                                  : DataType(DT::SFLOAT);
    }
    /* datatype.cpp:95-106 */
    static DataType SuggestComplex(DataType t) {
        return (t.dt_ == DT::UINT32 || t.dt_ == DT::SINT32 || t.dt_ == DT::DFLOAT || t.dt_ == DT::DCOMPLEX)
                   ? DataType(DT::DCOMPLEX)
                   : DataType(DT::SCOMPLEX);
    }

  private:
    DT dt_;
};
constexpr DataType DT_UINT8{DataType::DT::UINT8};
constexpr DataType DT_UINT16{DataType::DT::UINT16};
constexpr DataType DT_UINT32{DataType::DT::UINT32};
constexpr DataType DT_SINT8{DataType::DT::SINT8};
constexpr DataType DT_SINT16{DataType::DT::SINT16};
constexpr DataType DT_SINT32{DataType::DT::SINT32};
constexpr DataType DT_SFLOAT{DataType::DT::SFLOAT};
constexpr DataType DT_DFLOAT{DataType::DT::DFLOAT};
constexpr DataType DT_BIN{DataType::DT::BIN};
constexpr DataType DT_SCOMPLEX{DataType::DT::SCOMPLEX};
constexpr DataType DT_DCOMPLEX{DataType::DT::DCOMPLEX};

using DataSegment = std::shared_ptr<void>;
/* image.h:80-87: a shared pointer that does not own the data. */
inline DataSegment NonOwnedRefToDataSegment(void *ptr) { return DataSegment(ptr, [](void *) {}); }
inline DataSegment NonOwnedRefToDataSegment(void const *ptr) {
    return DataSegment(const_cast<void *>(ptr), [](void *) {});
}

namespace Option {
enum class AcceptDataTypeChange { DONT_ALLOW, DO_ALLOW };
}

/* Where an image's pixels live: device memory (the default, what the kernels read) or
 * host memory (the reference's ordinary images; every function copies such an image
 * to the device, runs there and copies the result back). */
enum class MemoryLocation { DEVICE, HOST };

/* The part of dip::Image the Gaussian path uses: sizes, strides in samples (dim 0
 * fastest), a tensor of one or more elements per pixel, data type, data segment and
 * its memory location.  Forge lays a tensor image out as planes (tensor stride = the
 * number of pixels), so every element is a contiguous image. */
class Image {
  public:
    Image() = default;
    /* Forged image of the given sizes, tensor elements and type (normal strides). */
    explicit Image(UnsignedArray sizes, uint tensorElems = 1, dip::DataType dt = DT_SFLOAT,
                   MemoryLocation where = MemoryLocation::DEVICE)
        : dataType_(dt), sizes_(std::move(sizes)), tensorElements_(tensorElems), onHost_(where == MemoryLocation::HOST) {
        if (tensorElems < 1) throw ParameterError("Tensor must have at least one element");
        Forge();
    }
    /* Image around existing data (image.h:458-491); no ownership if `data` says so.  A
     * tensor image gives its element count and tensor stride (0: one element). */
    Image(DataSegment const &data, void *origin, dip::DataType dt, UnsignedArray sizes, IntegerArray strides = {},
          uint tensorElems = 1, sint tensorStride = 0, MemoryLocation where = MemoryLocation::DEVICE)
        : dataType_(dt), sizes_(std::move(sizes)), strides_(std::move(strides)), tensorElements_(tensorElems),
          tensorStride_(tensorStride), onHost_(where == MemoryLocation::HOST), data_(data), origin_(origin) {
        if (!data || !origin) throw ParameterError("Bad data pointer");
        if (tensorElems < 1) throw ParameterError("Tensor must have at least one element");
        if (strides_.empty()) SetNormalStrides();
        if (strides_.size() != sizes_.size()) throw ParameterError("Strides array size does not match image dimensionality");
        if (tensorElems == 1) tensorStride_ = 1;
    }

    bool IsForged() const { return origin_ != nullptr; }
    uint Dimensionality() const { return sizes_.size(); }
    UnsignedArray const &Sizes() const { return sizes_; }
    uint Size(uint d) const { return sizes_.at(d); }
    IntegerArray const &Strides() const { return strides_; }
    sint Stride(uint d) const { return strides_.at(d); }
    uint TensorElements() const { return tensorElements_; }
    sint TensorStride() const { return tensorStride_; }
    bool IsScalar() const { return tensorElements_ == 1; }
    bool IsOnHost() const { return onHost_; }
    dip::DataType DataType() const { return dataType_; }
    void *Origin() const { return origin_; }
    DataSegment const &Data() const { return data_; }
    uint NumberOfPixels() const {
        uint n = 1;
        for (uint s : sizes_) n *= s;
        return n;
    }
    uint NumberOfSamples() const { return NumberOfPixels() * tensorElements_; }
    bool IsProtected() const { return protect_; }
    bool Protect(bool set = true) {
        bool old = protect_;
        protect_ = set;
        return old;
    }
    void SetSizes(UnsignedArray s) {
        if (IsForged()) throw ParameterError("Image is forged");
        sizes_ = std::move(s);
    }
    void SetDataType(dip::DataType dt) {
        if (IsForged()) throw ParameterError("Image is forged");
        dataType_ = dt;
    }
    void SetTensorElements(uint n) {
        if (IsForged()) throw ParameterError("Image is forged");
        if (n < 1) throw ParameterError("Tensor must have at least one element");
        tensorElements_ = n;
    }
    /* Memory of the next Forge (an unforged image only). */
    void SetMemoryLocation(MemoryLocation where) {
        if (IsForged()) throw ParameterError("Image is forged");
        onHost_ = where == MemoryLocation::HOST;
    }
    /* Tensor element k as a scalar image sharing the data. */
    Image TensorElement(uint k) const {
        if (!IsForged()) throw ParameterError("Image is not forged");
        if (k >= tensorElements_) throw ParameterError("Index out of range");
        void *o = static_cast<char *>(origin_) + static_cast<sint>(k) * tensorStride_ * static_cast<sint>(dataType_.SizeOf());
        Image e(data_, o, dataType_, sizes_, strides_, 1, 0, onHost_ ? MemoryLocation::HOST : MemoryLocation::DEVICE);
        e.protect_ = protect_;
        return e;
    }
    /* Allocates memory for the current sizes, tensor and type (normal strides; a
     * tensor as planes): device memory (hipMalloc) or pinned host memory. */
    void Forge() {
        if (IsForged()) return;
        SetNormalStrides();
        tensorStride_ = tensorElements_ > 1 ? static_cast<sint>(NumberOfPixels()) : 1;
        std::size_t bytes = NumberOfSamples() * dataType_.SizeOf();
        void *p = nullptr;
        if (onHost_) {
            if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess)
                throw RunTimeError("hipHostMalloc failed");
            data_ = DataSegment(p, [](void *q) { (void)hipHostFree(q); });
        } else {
            if (hipMalloc(&p, bytes ? bytes : 1) != hipSuccess) throw RunTimeError("hipMalloc failed");
            data_ = DataSegment(p, [](void *q) { (void)hipFree(q); });
        }
        origin_ = p;
    }
    /* image.h ReForge: keep the data when sizes, tensor and type fit, else reallocate
     * (or throw for a protected image; a protected image keeps its type). */
    void ReForge(UnsignedArray const &sizes, uint tensorElems, dip::DataType dt,
                 Option::AcceptDataTypeChange acceptDataTypeChange = Option::AcceptDataTypeChange::DONT_ALLOW) {
        if (IsForged()) {
            if (sizes_ == sizes && tensorElements_ == tensorElems &&
                (dataType_ == dt || (protect_ && acceptDataTypeChange == Option::AcceptDataTypeChange::DO_ALLOW)))
                return;
            if (protect_) throw ParameterError("Image is protected");
            Strip();
        }
        sizes_ = sizes;
        tensorElements_ = tensorElems;
        dataType_ = dt;
        Forge();
    }
    void Strip() {
        if (protect_) throw ParameterError("Image is protected");
        data_.reset();
        origin_ = nullptr;
        strides_.clear();
    }
    /* Normal strides and, for a tensor, planes: the samples form one contiguous block. */
    bool IsContiguousBlock() const {
        sint s = 1;
        for (uint d = 0; d < sizes_.size(); ++d) {
            if (sizes_[d] > 1 && strides_[d] != s) return false;
            s *= static_cast<sint>(sizes_[d]);
        }
        return tensorElements_ == 1 || tensorStride_ == s;
    }
    /* Copies between this image and a host buffer of NumberOfSamples() samples (the
     * planes of a tensor one after the other); the image's samples must form one block. */
    void CopyFromHost(void const *host) {
        if (!IsForged()) throw ParameterError("Image is not forged");
        if (!IsContiguousBlock()) throw ParameterError("Copy needs normal strides");
        if (hipMemcpy(origin_, host, NumberOfSamples() * dataType_.SizeOf(), hipMemcpyDefault) != hipSuccess)
            throw RunTimeError("hipMemcpy failed");
    }
    void CopyToHost(void *host) const {
        if (!IsForged()) throw ParameterError("Image is not forged");
        if (!IsContiguousBlock()) throw ParameterError("Copy needs normal strides");
        if (hipMemcpy(host, origin_, NumberOfSamples() * dataType_.SizeOf(), hipMemcpyDefault) != hipSuccess)
            throw RunTimeError("hipMemcpy failed");
    }
    /* The C ABI view of this image; a tensor image as a scalar image with the tensor as
     * one more dim (the reference's AsScalarImage, framework_separable.cpp:105-116). */
    dipamd_image View() const {
        dipamd_image v{};
        v.origin = origin_;
        v.dtype = dataType_.Code();
        v.ndim = static_cast<int32_t>(sizes_.size());
        for (uint d = 0; d < sizes_.size() && d < DIPAMD_MAX_DIMS; ++d) {
            v.sizes[d] = static_cast<int64_t>(sizes_[d]);
            v.strides[d] = static_cast<int64_t>(strides_[d]);
        }
        if (tensorElements_ > 1) {
            if (sizes_.size() >= DIPAMD_MAX_DIMS) throw ParameterError("Dimensionality not supported");
            v.sizes[v.ndim] = static_cast<int64_t>(tensorElements_);
            v.strides[v.ndim] = static_cast<int64_t>(tensorStride_);
            ++v.ndim;
        }
        return v;
    }

  private:
    void SetNormalStrides() {
        strides_.resize(sizes_.size());
        sint s = 1;
        for (uint d = 0; d < sizes_.size(); ++d) {
            strides_[d] = s;
            s *= static_cast<sint>(sizes_[d]);
        }
    }

    dip::DataType dataType_ = DT_SFLOAT;
    UnsignedArray sizes_;
    IntegerArray strides_;
    uint tensorElements_ = 1;
    sint tensorStride_ = 1;
    bool onHost_ = false;
    DataSegment data_;
    void *origin_ = nullptr;
    bool protect_ = false;
};

/* linear.h:84-90 */
struct OneDimensionalFilter {
    std::vector<dfloat> filter;
    sint origin = -1;
    String symmetry;
    bool isComplex = false;
};
using OneDimensionalFilterArray = std::vector<OneDimensionalFilter>;

namespace detail {
inline void Check(int status) {
    if (status == DIPAMD_OK) return;
    const char *m = dipamd_last_error();
    if (status == DIPAMD_E_HIP) throw RunTimeError(m);
    throw ParameterError(m);
}
inline int32_t BoundaryCode(String const &bc) {
    static const char *names[] = {"mirror", "asym mirror", "antisym reflect", "periodic", "asym periodic",
                                  "add zeros", "add max", "add min", "zero order", "first order",
                                  "second order", "third order", "already expanded"};
    if (bc.empty() || bc == "default") return DIPAMD_BC_SYMMETRIC_MIRROR;
    for (int32_t i = 0; i < 13; ++i)
        if (bc == names[i]) return i;
    throw ParameterError("Boundary condition not recognized: " + bc);
}
inline std::vector<int32_t> BoundaryCodes(StringArray const &bc) {
    std::vector<int32_t> out;
    for (auto const &b : bc) out.push_back(BoundaryCode(b));
    return out;
}
inline int32_t SymmetryCode(String const &s) {
    if (s.empty() || s == "general") return DIPAMD_SYM_GENERAL;
    if (s == "even") return DIPAMD_SYM_EVEN;
    if (s == "odd") return DIPAMD_SYM_ODD;
    if (s == "conj") return DIPAMD_SYM_CONJ;
    if (s == "d-even") return DIPAMD_SYM_D_EVEN;
    if (s == "d-odd") return DIPAMD_SYM_D_ODD;
    if (s == "d-conj") return DIPAMD_SYM_D_CONJ;
    throw ParameterError("Symmetry string not recognized: " + s);
}
/* ReForge `out` for a result of type `dt` from `in` (framework_separable.cpp:72-76, 126),
 * with the input's tensor. */
inline void PrepareOutput(Image const &in, Image &out, DataType dt) {
    if (!in.IsForged()) throw ParameterError("Image is not forged");
    const bool same = out.IsForged() && out.Origin() == in.Origin() && out.Sizes() == in.Sizes() &&
                      out.Strides() == in.Strides() && out.DataType() == in.DataType() &&
                      out.TensorElements() == in.TensorElements() && out.TensorStride() == in.TensorStride();
    if (same) {
        if (out.DataType() != dt && !out.IsProtected()) {
            // in-place but the output type changes: a new data segment for out
            Image fresh(in.Sizes(), in.TensorElements(), dt);
            out = fresh;
        }
        return;
    }
    out.ReForge(in.Sizes(), in.TensorElements(), dt, Option::AcceptDataTypeChange::DO_ALLOW);
}

/* Host images: the call `run(deviceIn, deviceOut)` works on a device copy of a host
 * input (normal strides needed) and, when `out` is in host memory (or unforged with a
 * host input), into a device image whose result is copied back into `out`.  Returns
 * false (and runs nothing) when every image is on the device. */
template <typename F>
bool Staged(Image const &in, Image &out, F &&run) {
    const bool outHost = out.IsForged() ? out.IsOnHost() : in.IsOnHost();
    if (!in.IsOnHost() && !outHost) return false;
    if (!in.IsForged()) throw ParameterError("Image is not forged");
    Image din = in;
    if (in.IsOnHost()) {
        if (!in.IsContiguousBlock()) throw ParameterError("Host images need normal strides");
        din = Image(in.Sizes(), in.TensorElements(), in.DataType());
        if (hipMemcpyAsync(din.Origin(), in.Origin(), in.NumberOfSamples() * in.DataType().SizeOf(), hipMemcpyDefault,
                           dipamd::CurrentStream()) != hipSuccess)
            throw RunTimeError("hipMemcpy failed");
    }
    if (!outHost) {
        run(din, out);
        if (hipStreamSynchronize(dipamd::CurrentStream()) != hipSuccess) throw RunTimeError("stream failed");
        return true;
    }
    Image dout;
    if (out.IsForged()) {
        dout = Image(out.Sizes(), out.TensorElements(), out.DataType());
        dout.Protect(out.IsProtected());
    }
    run(din, dout);
    if (hipStreamSynchronize(dipamd::CurrentStream()) != hipSuccess) throw RunTimeError("stream failed");
    if (!out.IsForged()) out.SetMemoryLocation(MemoryLocation::HOST);
    out.ReForge(dout.Sizes(), dout.TensorElements(), dout.DataType(), Option::AcceptDataTypeChange::DO_ALLOW);
    if (!out.IsContiguousBlock()) throw ParameterError("Host images need normal strides");
    if (hipMemcpy(out.Origin(), dout.Origin(), dout.NumberOfSamples() * dout.DataType().SizeOf(), hipMemcpyDefault) !=
        hipSuccess)
        throw RunTimeError("hipMemcpy failed");
    return true;
}

/* A per-dim parameter for the C ABI view of a tensor image, whose last dim is the
 * tensor: broadcast to the image's dims, then `fill` for the tensor dim (which no
 * filter processes).  Scalar images keep the parameter as given. */
template <typename T>
std::vector<T> TensorDim(std::vector<T> v, Image const &in, T dflt, T fill) {
    if (in.IsScalar()) return v;
    const uint nd = in.Dimensionality();
    if (v.empty()) {
        v.assign(nd, dflt);
    } else if (v.size() == 1) {
        v.assign(nd, v[0]);
    } else if (v.size() != nd) {
        throw ParameterError("Array parameter has the wrong number of elements");
    }
    v.push_back(fill);
    return v;
}
inline StringArray TensorDimBC(StringArray bc, Image const &in) {
    if (in.IsScalar() || bc.size() <= 1) return bc;  // broadcast by the C ABI
    return TensorDim<String>(std::move(bc), in, "", "mirror");
}
inline void PrepareOutput(Image const &in, Image &out) { PrepareOutput(in, out, DataType::SuggestFlex(in.DataType())); }
}  // namespace detail

/* dip::SeparableConvolution (linear.h:114-130); a complex filter (filter holds re, im
 * pairs) makes the result SuggestComplex(in), as convolution.cpp:331-346. */
inline void SeparableConvolution(Image const &in, Image &out, OneDimensionalFilterArray const &c_filterArray,
                                 StringArray const &c_boundaryCondition = {}, BooleanArray process = {}) {
    if (detail::Staged(in, out, [&](Image const &i, Image &o) {
            SeparableConvolution(i, o, c_filterArray, c_boundaryCondition, process);
        }))
        return;
    if (!in.IsForged()) throw ParameterError("Image is not forged");
    // a tensor image: its tensor dim gets the identity filter and is not processed
    OneDimensionalFilterArray filterArray =
        detail::TensorDim<OneDimensionalFilter>(c_filterArray, in, OneDimensionalFilter{}, OneDimensionalFilter{{1.0}});
    const StringArray boundaryCondition = detail::TensorDimBC(c_boundaryCondition, in);
    {
        std::vector<char> p(process.begin(), process.end());
        p = detail::TensorDim<char>(p, in, 1, 0);
        process.assign(p.begin(), p.end());
    }
    Image src = in;  // keeps the input's data alive if `out` is `in`
    bool complexFilter = false;
    for (auto const &o : filterArray) complexFilter = complexFilter || o.isComplex;
    detail::PrepareOutput(in, out, complexFilter ? DataType::SuggestComplex(in.DataType())
                                                 : DataType::SuggestFlex(in.DataType()));
    std::vector<dipamd_filter> f;
    for (auto const &o : filterArray) {
        if (o.isComplex && (o.filter.size() & 1)) throw ParameterError("Complex filter must have an even number of values");
        const size_t n = o.isComplex ? o.filter.size() / 2 : o.filter.size();
        f.push_back(dipamd_filter{o.filter.data(), static_cast<int32_t>(n), static_cast<int32_t>(o.origin),
                                  detail::SymmetryCode(o.symmetry), o.isComplex ? 1 : 0});
    }
    std::vector<int32_t> bc = detail::BoundaryCodes(boundaryCondition);
    std::vector<uint8_t> proc(process.begin(), process.end());
    dipamd_image a = src.View(), b = out.View();
    detail::Check(dipamd_separable_convolution(&a, &b, f.data(), static_cast<int32_t>(f.size()), bc.data(),
                                               static_cast<int32_t>(bc.size()), proc.empty() ? nullptr : proc.data(),
                                               DIPAMD_EXACT, dipamd::CurrentStream()));
}
inline Image SeparableConvolution(Image const &in, OneDimensionalFilterArray const &filter,
                                  StringArray const &boundaryCondition = {}, BooleanArray process = {}) {
    Image out;
    SeparableConvolution(in, out, filter, boundaryCondition, std::move(process));
    return out;
}

/* dip::GaussFIR (linear.h:300-318). */
inline void GaussFIR(Image const &in, Image &out, FloatArray sigmas = {1.0}, UnsignedArray derivativeOrder = {0},
                     StringArray const &boundaryCondition = {}, dfloat truncation = 3) {
    if (detail::Staged(in, out, [&](Image const &i, Image &o) {
            GaussFIR(i, o, sigmas, derivativeOrder, boundaryCondition, truncation);
        }))
        return;
    if (!in.IsForged()) throw ParameterError("Image is not forged");
    sigmas = detail::TensorDim<dfloat>(std::move(sigmas), in, 1.0, 0.0);
    derivativeOrder = detail::TensorDim<uint>(std::move(derivativeOrder), in, 0, 0);
    Image src = in;
    detail::PrepareOutput(in, out);
    std::vector<int32_t> ord(derivativeOrder.begin(), derivativeOrder.end());
    std::vector<int32_t> bc = detail::BoundaryCodes(detail::TensorDimBC(boundaryCondition, in));
    dipamd_image a = src.View(), b = out.View();
    detail::Check(dipamd_gauss_fir(&a, &b, sigmas.data(), static_cast<int32_t>(sigmas.size()), ord.data(),
                                   static_cast<int32_t>(ord.size()), bc.data(), static_cast<int32_t>(bc.size()),
                                   truncation, DIPAMD_EXACT, dipamd::CurrentStream()));
}
inline Image GaussFIR(Image const &in, FloatArray sigmas = {1.0}, UnsignedArray derivativeOrder = {0},
                      StringArray const &boundaryCondition = {}, dfloat truncation = 3) {
    Image out;
    GaussFIR(in, out, std::move(sigmas), std::move(derivativeOrder), boundaryCondition, truncation);
    return out;
}

/* dip::GaussIIR (linear.h:407-429). */
inline void GaussIIR(Image const &in, Image &out, FloatArray sigmas = {1.0}, UnsignedArray derivativeOrder = {0},
                     StringArray const &boundaryCondition = {}, UnsignedArray filterOrder = {},
                     String const &designMethod = "discrete time fit", dfloat truncation = 3) {
    if (detail::Staged(in, out, [&](Image const &i, Image &o) {
            GaussIIR(i, o, sigmas, derivativeOrder, boundaryCondition, filterOrder, designMethod, truncation);
        }))
        return;
    if (!in.IsForged()) throw ParameterError("Image is not forged");
    sigmas = detail::TensorDim<dfloat>(std::move(sigmas), in, 1.0, 0.0);
    derivativeOrder = detail::TensorDim<uint>(std::move(derivativeOrder), in, 0, 0);
    if (!filterOrder.empty()) filterOrder = detail::TensorDim<uint>(std::move(filterOrder), in, 3, 3);
    int32_t method;
    if (designMethod == "discrete time fit") {
        method = DIPAMD_IIR_DISCRETE_TIME_FIT;
    } else if (designMethod == "forward backward") {
        method = DIPAMD_IIR_FORWARD_BACKWARD;
    } else {
        throw ParameterError("Invalid flag: " + designMethod);
    }
    Image src = in;
    detail::PrepareOutput(in, out);
    std::vector<int32_t> ord(derivativeOrder.begin(), derivativeOrder.end());
    std::vector<int32_t> fo(filterOrder.begin(), filterOrder.end());
    std::vector<int32_t> bc = detail::BoundaryCodes(detail::TensorDimBC(boundaryCondition, in));
    dipamd_image a = src.View(), b = out.View();
    detail::Check(dipamd_gauss_iir(&a, &b, sigmas.data(), static_cast<int32_t>(sigmas.size()), ord.data(),
                                   static_cast<int32_t>(ord.size()), bc.data(), static_cast<int32_t>(bc.size()),
                                   fo.data(), static_cast<int32_t>(fo.size()), method, truncation, 0,
                                   dipamd::CurrentStream()));
}
inline Image GaussIIR(Image const &in, FloatArray sigmas = {1.0}, UnsignedArray derivativeOrder = {0},
                      StringArray const &boundaryCondition = {}, UnsignedArray filterOrder = {},
                      String const &designMethod = "discrete time fit", dfloat truncation = 3) {
    Image out;
    GaussIIR(in, out, std::move(sigmas), std::move(derivativeOrder), boundaryCondition, std::move(filterOrder),
             designMethod, truncation);
    return out;
}

/* dip::GaussFT (linear.h:355-364); "frequency" representations give / take the centred
 * spectrum (complex), without a boundary condition. */
inline void GaussFT(Image const &in, Image &out, FloatArray sigmas = {1.0}, UnsignedArray derivativeOrder = {0},
                    dfloat truncation = 3, String const &inRepresentation = "spatial",
                    String const &outRepresentation = "spatial", StringArray const &boundaryCondition = {}) {
    if (detail::Staged(in, out, [&](Image const &i, Image &o) {
            GaussFT(i, o, sigmas, derivativeOrder, truncation, inRepresentation, outRepresentation, boundaryCondition);
        }))
        return;
    if (!in.IsForged()) throw ParameterError("Image is not forged");
    sigmas = detail::TensorDim<dfloat>(std::move(sigmas), in, 1.0, 0.0);
    derivativeOrder = detail::TensorDim<uint>(std::move(derivativeOrder), in, 0, 0);
    for (auto const *r : {&inRepresentation, &outRepresentation})
        if (*r != "spatial" && *r != "frequency") throw ParameterError("Invalid flag: " + *r);
    const uint32_t reprFlags = (inRepresentation == "frequency" ? DIPAMD_FT_IN_FREQUENCY : 0u) |
                               (outRepresentation == "frequency" ? DIPAMD_FT_OUT_FREQUENCY : 0u);
    Image src = in;
    detail::PrepareOutput(in, out, reprFlags ? DataType::SuggestComplex(in.DataType())
                                             : DataType::SuggestFlex(in.DataType()));
    std::vector<int32_t> ord(derivativeOrder.begin(), derivativeOrder.end());
    std::vector<int32_t> bc = detail::BoundaryCodes(detail::TensorDimBC(boundaryCondition, in));
    dipamd_image a = src.View(), b = out.View();
    detail::Check(dipamd_gauss_ft(&a, &b, sigmas.data(), static_cast<int32_t>(sigmas.size()), ord.data(),
                                  static_cast<int32_t>(ord.size()), bc.data(), static_cast<int32_t>(bc.size()),
                                  truncation, reprFlags, dipamd::CurrentStream()));
}
inline Image GaussFT(Image const &in, FloatArray sigmas = {1.0}, UnsignedArray derivativeOrder = {0},
                     dfloat truncation = 3, String const &inRepresentation = "spatial",
                     String const &outRepresentation = "spatial", StringArray const &boundaryCondition = {}) {
    Image out;
    GaussFT(in, out, std::move(sigmas), std::move(derivativeOrder), truncation, inRepresentation, outRepresentation,
            boundaryCondition);
    return out;
}

/* dip::Gauss (linear.h:452-472; dispatch derivative.cpp:35-86): "FIR", "IIR", "FT" and
 * "best". */
inline void Gauss(Image const &in, Image &out, FloatArray sigmas = {1.0}, UnsignedArray derivativeOrder = {0},
                  String const &c_method = "best", StringArray const &boundaryCondition = {}, dfloat truncation = 3) {
    String method = c_method.substr(0, 5) == "gauss" ? c_method.substr(5) : c_method;
    if (method == "best") {
        method = "fir";
        for (auto o : derivativeOrder)
            if (o > 3) method = "ft";
        for (auto s : sigmas)
            if (s < 0.8 && s > 0.0) method = "ft";
        if (method == "fir")
            for (auto s : sigmas)
                if (s > 10) method = "iir";
    }
    if (method == "FIR" || method == "fir") {
        GaussFIR(in, out, std::move(sigmas), std::move(derivativeOrder), boundaryCondition, truncation);
        return;
    }
    if (method == "IIR" || method == "iir") {
        GaussIIR(in, out, std::move(sigmas), std::move(derivativeOrder), boundaryCondition, {}, "discrete time fit",
                 truncation);
        return;
    }
    if (method == "FT" || method == "ft") {
        // the empty condition is not GaussFT's default (derivative.cpp:74-77)
        GaussFT(in, out, std::move(sigmas), std::move(derivativeOrder), truncation, "spatial", "spatial",
                boundaryCondition.empty() ? StringArray{"default"} : boundaryCondition);
        return;
    }
    throw ParameterError("Invalid flag: " + c_method);
}
inline Image Gauss(Image const &in, FloatArray sigmas = {1.0}, UnsignedArray derivativeOrder = {0},
                   String const &method = "best", StringArray const &boundaryCondition = {}, dfloat truncation = 3) {
    Image out;
    Gauss(in, out, std::move(sigmas), std::move(derivativeOrder), method, boundaryCondition, truncation);
    return out;
}

/* The derivative family (derivative.cpp:120-521; linear.h:666, 697, 862, 904) with the
 * Gaussian FIR, IIR or FT method ("best" resolves to one of them), on scalar images.
 * Gradient, Hessian and StructureTensor give tensor images in dip's element order
 * (Hessian and the structure tensor: the diagonal first, then the upper triangle
 * column by column); overloads taking an ImageArray give the elements as separate
 * images. */
using ImageArray = std::vector<Image>;
namespace detail {
/* The Gaussian method the derivatives run (dip::Derivative passes `method` on to
 * dip::Gauss, derivative.cpp:111-114): DIPAMD_GAUSS_IIR for IIR, 0 for FIR. */
inline uint32_t GaussMethodFlag(String const &c_method, FloatArray const &sigmas) {
    String method = c_method == "gauss" ? "best" : (c_method.substr(0, 5) == "gauss" ? c_method.substr(5) : c_method);
    if (method == "best") {  // derivative orders <= 2 here: sigma decides (derivative.cpp:35-58)
        method = "fir";
        for (auto s : sigmas)
            if (s < 0.8 && s > 0.0) method = "ft";
        if (method == "fir")
            for (auto s : sigmas)
                if (s > 10) method = "iir";
    }
    if (method == "FIR" || method == "fir") return 0;
    if (method == "IIR" || method == "iir") return DIPAMD_GAUSS_IIR;
    if (method == "FT" || method == "ft") return DIPAMD_GAUSS_FT;
    if (method == "finitediff")
        throw ParameterError("method \"finitediff\" is not provided by this library (Gaussian methods only)");
    throw ParameterError("Invalid flag: " + c_method);
}
inline void RunDerivative(int32_t fn, Image const &in, std::vector<dipamd_image> &outs, FloatArray const &sigmas,
                          StringArray const &boundaryCondition, BooleanArray const &process, dfloat truncation,
                          uint32_t method_flag) {
    std::vector<int32_t> bc = BoundaryCodes(boundaryCondition);
    std::vector<uint8_t> proc(process.begin(), process.end());
    dipamd_image a = in.View();
    Check(dipamd_derivative(fn, &a, outs.data(), static_cast<int32_t>(outs.size()), sigmas.data(),
                            static_cast<int32_t>(sigmas.size()), bc.data(), static_cast<int32_t>(bc.size()),
                            proc.empty() ? nullptr : proc.data(), static_cast<int32_t>(proc.size()), truncation,
                            DIPAMD_EXACT | method_flag, dipamd::CurrentStream()));
}
inline int DifferentiatedDims(Image const &in, FloatArray const &sigmas, BooleanArray const &process) {
    const uint nd = in.Dimensionality();
    int n = 0;
    for (uint d = 0; d < nd; ++d) {
        const dfloat s = sigmas.empty() ? 1.0 : (sigmas.size() == 1 ? sigmas[0] : sigmas[d]);
        const bool p = process.empty() ? true : (process.size() == 1 ? process[0] : process[d]);
        if (p && in.Sizes()[d] > 1 && s > 0.0) ++n;
    }
    if (n < 1) throw ParameterError("Dimensionality not supported");
    return n;
}
inline void RequireScalar(Image const &in) {
    if (!in.IsForged()) throw ParameterError("Image is not forged");
    if (!in.IsScalar()) throw ParameterError("Image is not scalar");
}
/* The tensor result as one tensor image (each element a plane of `out`). */
inline void TensorResult(int32_t fn, Image const &in, Image &out, FloatArray const &sigmas, String const &method,
                         StringArray const &bc, BooleanArray const &process, dfloat truncation) {
    if (Staged(in, out, [&](Image const &i, Image &o) {
            TensorResult(fn, i, o, sigmas, method, bc, process, truncation);
        }))
        return;
    RequireScalar(in);
    const uint32_t mf = GaussMethodFlag(method, sigmas);
    const int n = dipamd_derivative_outputs(fn, DifferentiatedDims(in, sigmas, process));
    if (n < 0) Check(n);
    Image src = in;
    if (out.IsForged() && out.Origin() == in.Origin()) out = Image();  // the result never overwrites the input
    out.ReForge(in.Sizes(), static_cast<uint>(n), DataType::SuggestFlex(in.DataType()),
                Option::AcceptDataTypeChange::DO_ALLOW);
    std::vector<dipamd_image> views;
    for (int k = 0; k < n; ++k) views.push_back(out.TensorElement(static_cast<uint>(k)).View());
    RunDerivative(fn, src, views, sigmas, bc, process, truncation, mf);
}
inline void TensorResult(int32_t fn, Image const &in, ImageArray &out, FloatArray const &sigmas,
                         String const &method, StringArray const &bc, BooleanArray const &process,
                         dfloat truncation) {
    RequireScalar(in);
    if (in.IsOnHost()) {
        Image t;
        TensorResult(fn, in, t, sigmas, method, bc, process, truncation);
        out.clear();
        for (uint k = 0; k < t.TensorElements(); ++k) out.push_back(t.TensorElement(k));
        return;
    }
    const uint32_t mf = GaussMethodFlag(method, sigmas);
    const int n = dipamd_derivative_outputs(fn, DifferentiatedDims(in, sigmas, process));
    if (n < 0) Check(n);
    out.clear();
    std::vector<dipamd_image> views;
    for (int k = 0; k < n; ++k) {
        out.emplace_back(in.Sizes(), 1, DataType::SuggestFlex(in.DataType()));
        views.push_back(out.back().View());
    }
    RunDerivative(fn, in, views, sigmas, bc, process, truncation, mf);
}
inline void ScalarResult(int32_t fn, Image const &in, Image &out, FloatArray const &sigmas, String const &method,
                         StringArray const &bc, BooleanArray const &process, dfloat truncation) {
    if (Staged(in, out, [&](Image const &i, Image &o) {
            ScalarResult(fn, i, o, sigmas, method, bc, process, truncation);
        }))
        return;
    RequireScalar(in);
    const uint32_t mf = GaussMethodFlag(method, sigmas);
    DifferentiatedDims(in, sigmas, process);
    Image src = in;
    PrepareOutput(in, out);
    std::vector<dipamd_image> views{out.View()};
    RunDerivative(fn, src, views, sigmas, bc, process, truncation, mf);
}
}  // namespace detail

inline void Gradient(Image const &in, Image &out, FloatArray sigmas = {1.0}, String const &method = "best",
                     StringArray const &boundaryCondition = {}, BooleanArray process = {}, dfloat truncation = 3) {
    detail::TensorResult(DIPAMD_FN_GRADIENT, in, out, sigmas, method, boundaryCondition, process, truncation);
}
inline Image Gradient(Image const &in, FloatArray sigmas = {1.0}, String const &method = "best",
                      StringArray const &boundaryCondition = {}, BooleanArray process = {}, dfloat truncation = 3) {
    Image out;
    Gradient(in, out, std::move(sigmas), method, boundaryCondition, std::move(process), truncation);
    return out;
}
inline void Gradient(Image const &in, ImageArray &out, FloatArray sigmas = {1.0}, String const &method = "best",
                     StringArray const &boundaryCondition = {}, BooleanArray process = {}, dfloat truncation = 3) {
    detail::TensorResult(DIPAMD_FN_GRADIENT, in, out, sigmas, method, boundaryCondition, process, truncation);
}
inline void Hessian(Image const &in, Image &out, FloatArray sigmas = {1.0}, String const &method = "best",
                    StringArray const &boundaryCondition = {}, BooleanArray process = {}, dfloat truncation = 3) {
    detail::TensorResult(DIPAMD_FN_HESSIAN, in, out, sigmas, method, boundaryCondition, process, truncation);
}
inline Image Hessian(Image const &in, FloatArray sigmas = {1.0}, String const &method = "best",
                     StringArray const &boundaryCondition = {}, BooleanArray process = {}, dfloat truncation = 3) {
    Image out;
    Hessian(in, out, std::move(sigmas), method, boundaryCondition, std::move(process), truncation);
    return out;
}
inline void Hessian(Image const &in, ImageArray &out, FloatArray sigmas = {1.0}, String const &method = "best",
                    StringArray const &boundaryCondition = {}, BooleanArray process = {}, dfloat truncation = 3) {
    detail::TensorResult(DIPAMD_FN_HESSIAN, in, out, sigmas, method, boundaryCondition, process, truncation);
}
inline void GradientMagnitude(Image const &in, Image &out, FloatArray sigmas = {1.0}, String const &method = "best",
                              StringArray const &boundaryCondition = {}, BooleanArray process = {},
                              dfloat truncation = 3) {
    detail::ScalarResult(DIPAMD_FN_GRADIENT_MAGNITUDE, in, out, sigmas, method, boundaryCondition, process,
                         truncation);
}
inline Image GradientMagnitude(Image const &in, FloatArray sigmas = {1.0}, String const &method = "best",
                               StringArray const &boundaryCondition = {}, BooleanArray process = {},
                               dfloat truncation = 3) {
    Image out;
    GradientMagnitude(in, out, std::move(sigmas), method, boundaryCondition, std::move(process), truncation);
    return out;
}
inline void Laplace(Image const &in, Image &out, FloatArray sigmas = {1.0}, String const &method = "best",
                    StringArray const &boundaryCondition = {}, BooleanArray process = {}, dfloat truncation = 3) {
    detail::ScalarResult(DIPAMD_FN_LAPLACE, in, out, sigmas, method, boundaryCondition, process, truncation);
}
inline Image Laplace(Image const &in, FloatArray sigmas = {1.0}, String const &method = "best",
                     StringArray const &boundaryCondition = {}, BooleanArray process = {}, dfloat truncation = 3) {
    Image out;
    Laplace(in, out, std::move(sigmas), method, boundaryCondition, std::move(process), truncation);
    return out;
}
/* dip::Dgg, LaplacePlusDgg, LaplaceMinusDgg (derivative.cpp:486-520). */
inline void Dgg(Image const &in, Image &out, FloatArray sigmas = {1.0}, String const &method = "best",
                StringArray const &boundaryCondition = {}, BooleanArray process = {}, dfloat truncation = 3) {
    detail::ScalarResult(DIPAMD_FN_DGG, in, out, sigmas, method, boundaryCondition, process, truncation);
}
inline void LaplacePlusDgg(Image const &in, Image &out, FloatArray sigmas = {1.0}, String const &method = "best",
                           StringArray const &boundaryCondition = {}, BooleanArray process = {},
                           dfloat truncation = 3) {
    detail::ScalarResult(DIPAMD_FN_LAPLACE_PLUS_DGG, in, out, sigmas, method, boundaryCondition, process, truncation);
}
inline void LaplaceMinusDgg(Image const &in, Image &out, FloatArray sigmas = {1.0}, String const &method = "best",
                            StringArray const &boundaryCondition = {}, BooleanArray process = {},
                            dfloat truncation = 3) {
    detail::ScalarResult(DIPAMD_FN_LAPLACE_MINUS_DGG, in, out, sigmas, method, boundaryCondition, process,
                         truncation);
}
/* dip::StructureTensor (analysis.h:454; structure_tensor.cpp:32-57), with or without a
 * mask: the symmetric-matrix elements as a tensor image (or, with an ImageArray, as
 * separate images). */
inline void StructureTensor(Image const &in, Image const &mask, Image &out, FloatArray gradientSigmas = {1.0},
                            FloatArray tensorSigmas = {5.0}, String const &method = "best",
                            StringArray const &boundaryCondition = {}, dfloat truncation = 3) {
    if (detail::Staged(in, out, [&](Image const &i, Image &o) {
            Image m = mask;
            if (mask.IsForged() && mask.IsOnHost()) {  // a device copy of a host mask
                m = Image(mask.Sizes(), 1, mask.DataType());
                m.CopyFromHost(mask.Origin());
            }
            StructureTensor(i, m, o, gradientSigmas, tensorSigmas, method, boundaryCondition, truncation);
        }))
        return;
    detail::RequireScalar(in);
    const uint32_t tf = detail::GaussMethodFlag(method, tensorSigmas);
    const uint32_t flags = detail::GaussMethodFlag(method, gradientSigmas) |
                           (tf == DIPAMD_GAUSS_IIR ? DIPAMD_TENSOR_IIR : (tf == DIPAMD_GAUSS_FT ? DIPAMD_TENSOR_FT : 0u));
    const int n = dipamd_derivative_outputs(DIPAMD_FN_STRUCTURE_TENSOR,
                                            detail::DifferentiatedDims(in, gradientSigmas, {}));
    if (n < 0) detail::Check(n);
    Image src = in;
    if (out.IsForged() && out.Origin() == in.Origin()) out = Image();
    out.ReForge(in.Sizes(), static_cast<uint>(n), DataType::SuggestFlex(in.DataType()),
                Option::AcceptDataTypeChange::DO_ALLOW);
    std::vector<dipamd_image> views;
    std::vector<int32_t> bc = detail::BoundaryCodes(boundaryCondition);
    dipamd_image a = src.View();
    if (mask.IsForged()) {
        // structure_tensor.cpp:46-52: the normalized derivative of every dim
        const uint nd = in.Dimensionality();
        out.ReForge(in.Sizes(), nd * (nd + 1) / 2, DataType::SuggestFlex(in.DataType()),
                    Option::AcceptDataTypeChange::DO_ALLOW);
        for (uint k = 0; k < out.TensorElements(); ++k) views.push_back(out.TensorElement(k).View());
        dipamd_image m = mask.View();
        detail::Check(dipamd_structure_tensor_masked(&a, &m, views.data(), static_cast<int32_t>(views.size()),
                                                     gradientSigmas.data(), static_cast<int32_t>(gradientSigmas.size()),
                                                     tensorSigmas.data(), static_cast<int32_t>(tensorSigmas.size()),
                                                     bc.data(), static_cast<int32_t>(bc.size()), truncation,
                                                     DIPAMD_EXACT | flags, dipamd::CurrentStream()));
        return;
    }
    for (int k = 0; k < n; ++k) views.push_back(out.TensorElement(static_cast<uint>(k)).View());
    detail::Check(dipamd_structure_tensor(&a, views.data(), n, gradientSigmas.data(),
                                          static_cast<int32_t>(gradientSigmas.size()), tensorSigmas.data(),
                                          static_cast<int32_t>(tensorSigmas.size()), bc.data(),
                                          static_cast<int32_t>(bc.size()), truncation, DIPAMD_EXACT | flags,
                                          dipamd::CurrentStream()));
}
inline void StructureTensor(Image const &in, Image const &mask, ImageArray &out, FloatArray gradientSigmas = {1.0},
                            FloatArray tensorSigmas = {5.0}, String const &method = "best",
                            StringArray const &boundaryCondition = {}, dfloat truncation = 3) {
    Image t;
    StructureTensor(in, mask, t, std::move(gradientSigmas), std::move(tensorSigmas), method, boundaryCondition,
                    truncation);
    out.clear();
    for (uint k = 0; k < t.TensorElements(); ++k) out.push_back(t.TensorElement(k));
}

/* dip::NormalizedConvolution and dip::NormalizedDifferentialConvolution
 * (derivative.cpp:522-600) with the Gaussian methods. */
inline void NormalizedConvolution(Image const &in, Image const &mask, Image &out, FloatArray sigmas = {1.0},
                                  String const &method = "best", StringArray const &boundaryCondition = {},
                                  dfloat truncation = 3) {
    detail::RequireScalar(in);
    if (in.IsOnHost() || mask.IsOnHost() || (out.IsForged() && out.IsOnHost()))
        throw ParameterError("NormalizedConvolution takes device images");
    Image src = in;
    detail::PrepareOutput(in, out);
    std::vector<int32_t> bc = detail::BoundaryCodes(boundaryCondition);
    dipamd_image a = src.View(), m = mask.View(), b = out.View();
    detail::Check(dipamd_normalized_convolution(&a, &m, &b, sigmas.data(), static_cast<int32_t>(sigmas.size()),
                                                bc.data(), static_cast<int32_t>(bc.size()), truncation,
                                                detail::GaussMethodFlag(method, sigmas), dipamd::CurrentStream()));
}
inline void NormalizedDifferentialConvolution(Image const &in, Image const &mask, Image &out, uint dimension = 0,
                                              FloatArray sigmas = {1.0}, String const &method = "best",
                                              StringArray const &boundaryCondition = {}, dfloat truncation = 3) {
    detail::RequireScalar(in);
    if (in.IsOnHost() || mask.IsOnHost() || (out.IsForged() && out.IsOnHost()))
        throw ParameterError("NormalizedDifferentialConvolution takes device images");
    Image src = in;
    detail::PrepareOutput(in, out);
    std::vector<int32_t> bc = detail::BoundaryCodes(boundaryCondition);
    dipamd_image a = src.View(), m = mask.View(), b = out.View();
    detail::Check(dipamd_normalized_differential_convolution(
        &a, &m, &b, static_cast<int32_t>(dimension), sigmas.data(), static_cast<int32_t>(sigmas.size()), bc.data(),
        static_cast<int32_t>(bc.size()), truncation, detail::GaussMethodFlag(method, sigmas), dipamd::CurrentStream()));
}

/* dip::Derivative (derivative.cpp:88-115), Gaussian methods: dip::Gauss. */
inline void Derivative(Image const &in, Image &out, UnsignedArray derivativeOrder, FloatArray sigmas = {1.0},
                       String const &method = "best", StringArray const &boundaryCondition = {},
                       dfloat truncation = 3) {
    Gauss(in, out, std::move(sigmas), std::move(derivativeOrder), method, boundaryCondition, truncation);
}

/* dip::MakeHalfGaussian (generation.h:638-653; gauss.cpp:151-165). */
inline std::vector<dfloat> MakeHalfGaussian(dfloat sigma, uint derivativeOrder = 0, dfloat truncation = 3,
                                            DataType dt = DT_DFLOAT) {
    std::vector<dfloat> f(4096);
    const int n = dipamd_make_half_gaussian(sigma, static_cast<int32_t>(derivativeOrder), truncation,
                                            dt == DT_DFLOAT ? 1 : 0, f.data(), static_cast<int32_t>(f.size()));
    if (n < 0) detail::Check(n);
    f.resize(static_cast<std::size_t>(n));
    return f;
}

/* dip::SetNumberOfThreads / GetNumberOfThreads (multithreading.h:65-74;
 * multithreading.cpp:28-44).  The launch grid replaces the OpenMP team, so the value is
 * only kept (per host thread, default = the host's hardware threads, 0 resets it, larger
 * values are capped) and returned, as the reference keeps it. */
namespace detail {
inline uint DefaultThreadCount() {
    const unsigned n = std::thread::hardware_concurrency();
    return n ? static_cast<uint>(n) : 1;
}
inline uint &ThreadCount() {
    static thread_local uint n = DefaultThreadCount();
    return n;
}
}  // namespace detail
inline void SetNumberOfThreads(uint nThreads) {
    detail::ThreadCount() = nThreads == 0 ? detail::DefaultThreadCount()
                                          : std::min(nThreads, detail::DefaultThreadCount());
}
inline uint GetNumberOfThreads() { return detail::ThreadCount(); }

}  // namespace dip

#endif  // DIPLIB_AMD_HPP
