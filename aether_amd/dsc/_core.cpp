// _core.cpp -- aether_amd.dsc._core: the array type of the reference's Python package `aether`
// (dsc/src/aether_core.cpp:1-626, dsc/aether/__init__.py:1-161) on MI355X.
//
// One buffer on one device: host memory ("cpu"), page-locked host memory ("pinned") or HIP device memory
// ("cuda", the name torch uses for HIP devices too). Host and pinned arrays speak the buffer protocol;
// every array speaks DLPack (versioned and legacy capsules, deleters run exactly once); device arrays also
// expose __cuda_array_interface__. Device and pinned memory, fills, copies, stream ordering and FFT plans
// go through this library's `aether-memory/1` seam (libaether_hip.so, opened on the first device call),
// the table the reference's module binds (dsc/src/seam.h:16-51). Plain CPython C API: no torch, no numpy,
// no HIP headers.
#define PY_SSIZE_T_CLEAN
#include <Python.h>

#include <dlfcn.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/aether_hip.h"

namespace {

// ------------------------------------------------------------------ DLPack 1.0 (the public ABI)
struct DLDevice {
    int32_t device_type;
    int32_t device_id;
};
struct DLDataType {
    uint8_t code;
    uint8_t bits;
    uint16_t lanes;
};
struct DLTensor {
    void* data;
    DLDevice device;
    int32_t ndim;
    DLDataType dtype;
    int64_t* shape;
    int64_t* strides;
    uint64_t byte_offset;
};
struct DLManagedTensor {
    DLTensor dl_tensor;
    void* manager_ctx;
    void (*deleter)(DLManagedTensor*);
};
struct DLPackVersion {
    uint32_t major;
    uint32_t minor;
};
struct DLManagedTensorVersioned {
    DLPackVersion version;
    void* manager_ctx;
    void (*deleter)(DLManagedTensorVersioned*);
    uint64_t flags;
    DLTensor dl_tensor;
};
constexpr int32_t kDLCPU = 1, kDLCUDAHost = 3, kDLROCM = 10, kDLROCMHost = 11;
constexpr uint8_t kDLInt = 0, kDLUInt = 1, kDLFloat = 2, kDLBfloat = 4, kDLBool = 6;

constexpr int kRank = AETH_RUNTIME_RANK_CAP;
enum Kind { CPU = AETHER_MEM_CPU, DEV = AETHER_MEM_DEVICE, PINNED = AETHER_MEM_PINNED };

// ------------------------------------------------------------------ dtypes
struct TypeInfo {
    const char* name;
    uint8_t code, bits;
    const char* format;  // buffer-protocol format, null where Python has none (bfloat16)
    const char* typestr; // __cuda_array_interface__ typestr, null where none exists
};
enum : int { F64, F32, F16, BF16, I64, I32, I16, I8, U64, U32, U16, U8, BOOL, kTypes };
const TypeInfo kType[kTypes] = {
    {"float64", kDLFloat, 64, "d", "<f8"}, {"float32", kDLFloat, 32, "f", "<f4"},
    {"float16", kDLFloat, 16, "e", "<f2"}, {"bfloat16", kDLBfloat, 16, nullptr, nullptr},
    {"int64", kDLInt, 64, "q", "<i8"},     {"int32", kDLInt, 32, "i", "<i4"},
    {"int16", kDLInt, 16, "h", "<i2"},     {"int8", kDLInt, 8, "b", "|i1"},
    {"uint64", kDLUInt, 64, "Q", "<u8"},   {"uint32", kDLUInt, 32, "I", "<u4"},
    {"uint16", kDLUInt, 16, "H", "<u2"},   {"uint8", kDLUInt, 8, "B", "|u1"},
    {"bool", kDLBool, 8, "?", "|b1"},
};
size_t item(int t) { return kType[t].bits / 8; }

PyObject* g_error = nullptr;           // aether.Error
PyObject* g_unavailable = nullptr;     // aether.CudaUnavailable
PyObject* g_fft_unavailable = nullptr; // aether.FftUnavailable
PyObject* g_device_cls = nullptr;      // aether.Device
PyObject* g_dtype_cls = nullptr;       // aether.DType
std::atomic<long> g_exports{0};        // DLPack exports whose deleter has not run yet

// A Python error is set and the caller returns its failure value.
struct PyErrorSet {};

[[noreturn]] void fail(const std::string& m)
{
    PyErr_SetString(g_error, m.c_str());
    throw PyErrorSet{};
}
[[noreturn]] void fail_as(PyObject* type, const std::string& m)
{
    PyErr_SetString(type, m.c_str());
    throw PyErrorSet{};
}
[[noreturn]] void propagate() { throw PyErrorSet{}; }

int type_by_name(const char* name)
{
    for (int t = 0; t < kTypes; ++t)
        if (std::strcmp(name, kType[t].name) == 0) return t;
    fail(std::string("dtype '") + name +
         "' is not supported (float64 float32 float16 bfloat16 int64 int32 int16 int8 uint64 uint32 uint16 uint8 bool)");
}

int type_by_dl(DLDataType d)
{
    for (int t = 0; t < kTypes; ++t)
        if (kType[t].code == d.code && kType[t].bits == d.bits && d.lanes == 1) return t;
    fail("DLPack dtype (code " + std::to_string(d.code) + ", bits " + std::to_string(d.bits) + ", lanes " +
         std::to_string(d.lanes) + ") is not supported");
}

int type_by_format(const char* f)
{
    while (*f == '@' || *f == '=' || *f == '<' || *f == '|') ++f;
    std::string c(f);
    if (c == "l") c = "q"; // LP64: long is 8 bytes
    if (c == "L") c = "Q";
    for (int t = 0; t < kTypes; ++t)
        if (kType[t].format && c == kType[t].format) return t;
    fail("buffer format '" + std::string(f) + "' is not supported");
}

// float -> IEEE binary16, round to nearest even; subnormal results kept, overflow to infinity
uint16_t half_from_float(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    const uint16_t sign = (uint16_t)((x >> 16) & 0x8000u);
    const uint32_t ex = (x >> 23) & 0xffu, man = x & 0x7fffffu;
    if (ex == 0xffu) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u : 0u));
    if (ex == 0) return sign; // float subnormals lie far below half's smallest subnormal
    int e = (int)ex - 127;
    if (e > 15) return (uint16_t)(sign | 0x7c00u);
    const uint32_t sig = man | 0x800000u;               // 24 significant bits, value sig * 2^(e - 23)
    const int shift = 13 + ((e < -14 ? -14 : e) - e);   // bits below half's last place
    if (shift > 31) return sign;
    uint32_t q = sig >> shift;
    const uint32_t rem = sig & ((1u << shift) - 1u), halfway = 1u << (shift - 1);
    if (rem > halfway || (rem == halfway && (q & 1u))) ++q;
    if (e < -14) return (uint16_t)(sign | q);             // q == 1024 encodes the smallest normal
    if (q == 2048u) {
        q = 1024u;
        if (++e > 15) return (uint16_t)(sign | 0x7c00u);
    }
    return (uint16_t)(sign | (uint32_t)(e + 15) << 10 | (q - 1024u));
}

// float -> bfloat16, round to nearest even (quiet NaN kept quiet)
uint16_t bfloat_from_float(float f)
{
    uint32_t x;
    std::memcpy(&x, &f, 4);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    return (uint16_t)((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
}

// One element of dtype t from a Python scalar. 16-bit floats round through float first; integers take
// the low bits of the value as a C cast does.
void pack(int t, PyObject* v, unsigned char* out)
{
    auto as_double = [&] {
        const double d = PyFloat_AsDouble(v);
        if (d == -1.0 && PyErr_Occurred()) propagate();
        return d;
    };
    auto as_signed = [&] {
        const long long z = PyLong_AsLongLong(v);
        if (z == -1 && PyErr_Occurred()) propagate();
        return z;
    };
    auto as_unsigned = [&] {
        PyObject* n = PyNumber_Index(v);
        if (!n) propagate();
        const unsigned long long z = PyLong_AsUnsignedLongLongMask(n);
        Py_DECREF(n);
        if (z == (unsigned long long)-1 && PyErr_Occurred()) propagate();
        return z;
    };
    switch (t) {
    case F64: { const double d = as_double(); std::memcpy(out, &d, 8); break; }
    case F32: { const float f = (float)as_double(); std::memcpy(out, &f, 4); break; }
    case F16: { const uint16_t h = half_from_float((float)as_double()); std::memcpy(out, &h, 2); break; }
    case BF16: { const uint16_t h = bfloat_from_float((float)as_double()); std::memcpy(out, &h, 2); break; }
    case I64: { const int64_t z = as_signed(); std::memcpy(out, &z, 8); break; }
    case I32: { const int32_t z = (int32_t)as_signed(); std::memcpy(out, &z, 4); break; }
    case I16: { const int16_t z = (int16_t)as_signed(); std::memcpy(out, &z, 2); break; }
    case I8: { const int8_t z = (int8_t)as_signed(); std::memcpy(out, &z, 1); break; }
    case U64: { const uint64_t z = as_unsigned(); std::memcpy(out, &z, 8); break; }
    case U32: { const uint32_t z = (uint32_t)as_unsigned(); std::memcpy(out, &z, 4); break; }
    case U16: { const uint16_t z = (uint16_t)as_unsigned(); std::memcpy(out, &z, 2); break; }
    case U8: { out[0] = (unsigned char)as_unsigned(); break; }
    default: {
        const int b = PyObject_IsTrue(v);
        if (b < 0) propagate();
        out[0] = (unsigned char)b;
    }
    }
}

// ------------------------------------------------------------------ the seam (libaether_hip.so)
const AetherMemoryV1* g_seam = nullptr;
std::string g_seam_failure;
std::once_flag g_seam_once;

std::string module_dir()
{
    Dl_info info;
    if (!dladdr(reinterpret_cast<void*>(&module_dir), &info) || !info.dli_fname) return ".";
    std::string p = info.dli_fname;
    const size_t cut = p.find_last_of('/');
    return cut == std::string::npos ? "." : p.substr(0, cut);
}

void open_seam()
{
    const std::string path = module_dir() + "/../lib/libaether_hip.so";
    void* h = dlopen(path.c_str(), RTLD_NOW | RTLD_LOCAL);
    if (!h) {
        const char* e = dlerror();
        g_seam_failure = std::string("cannot load ") + path + ": " + (e ? e : "unknown error");
        return;
    }
    auto fn = reinterpret_cast<const AetherMemoryV1* (*)(void)>(dlsym(h, "aether_memory_v1"));
    const AetherMemoryV1* t = fn ? fn() : nullptr;
    if (!t || t->abi != AETHER_MEMORY_ABI) {
        g_seam_failure = path + " is not an aether-memory/" + std::to_string(AETHER_MEMORY_ABI) + " library";
        return;
    }
    int n = 0;
    if (t->probe(&n) != 0) {
        g_seam_failure = std::string("HIP runtime: ") + t->last_error();
        return;
    }
    g_seam = t;
}

// The seam, loaded on first use; CudaUnavailable when there is no usable HIP runtime or GPU.
const AetherMemoryV1* seam()
{
    if (std::getenv("AETHER_CUDA_MAJOR"))
        fail("AETHER_CUDA_MAJOR is no longer used: this library serves every supported runtime; unset it");
    std::call_once(g_seam_once, open_seam);
    if (!g_seam) fail_as(g_unavailable, "no GPU runtime available: " + g_seam_failure);
    return g_seam;
}

void check(int code, const std::string& what)
{
    if (code != 0) fail(what + ": " + seam()->last_error());
}

// ------------------------------------------------------------------ storage
// The memory behind one or more arrays and exports; released when the last holder goes. Whatever it
// releases that belongs to Python is released under the GIL.
struct Storage {
    void* ptr = nullptr;
    size_t bytes = 0;
    int kind = CPU, dev = 0;
    void* stream = nullptr;          // the allocation stream (device and pinned arrays)
    PyObject* stream_obj = nullptr;  // the caller's stream object: it outlives the array
    PyObject* keep = nullptr;        // a foreign object whose memory this borrows
    Py_buffer* view = nullptr;       // a borrowed buffer
    void* host = nullptr;            // this module's host allocation
    void* chunk = nullptr;           // a seam chunk (device or pinned)
    DLManagedTensorVersioned* imported_v = nullptr;
    DLManagedTensor* imported = nullptr;

    ~Storage()
    {
        if (chunk && g_seam) g_seam->chunk_release(chunk); // queued on its stream, deferred by a capture
        std::free(host);
        const bool py = Py_IsInitialized() != 0;
        PyGILState_STATE gil{};
        if (py) gil = PyGILState_Ensure();
        if (imported_v && imported_v->deleter) imported_v->deleter(imported_v);
        if (imported && imported->deleter) imported->deleter(imported);
        if (py) {
            if (view) {
                PyBuffer_Release(view);
                delete view;
            }
            Py_XDECREF(keep);
            Py_XDECREF(stream_obj);
            PyGILState_Release(gil);
        }
    }
};
using StoragePtr = std::shared_ptr<Storage>;

// ------------------------------------------------------------------ the Array type
struct ArrayObj {
    PyObject_HEAD
    StoragePtr st;
    int dt;
    int rank;
    int64_t ext[kRank];
    int64_t str[kRank]; // in elements
};
PyTypeObject ArrayType = {PyVarObject_HEAD_INIT(nullptr, 0)};

ArrayObj* as_array(PyObject* o, const char* what)
{
    if (!PyObject_TypeCheck(o, &ArrayType)) fail(std::string(what) + " is not an aether array");
    return reinterpret_cast<ArrayObj*>(o);
}

int64_t count(const ArrayObj* a)
{
    int64_t n = 1;
    for (int k = 0; k < a->rank; ++k) n *= a->ext[k];
    return n;
}
size_t nbytes(const ArrayObj* a) { return (size_t)count(a) * item(a->dt); }
bool contiguous(const ArrayObj* a)
{
    int64_t want = 1;
    for (int k = a->rank - 1; k >= 0; --k) {
        if (a->ext[k] != 1 && a->str[k] != want) return false;
        want *= a->ext[k];
    }
    return true;
}
bool host_side(const ArrayObj* a) { return a->st->kind != DEV; }
void need_contiguous(const ArrayObj* a, const char* what)
{
    if (!contiguous(a)) fail(std::string(what) + " needs a contiguous array; this one is strided");
}

// Extents that are valid for an array of `itemsize`-byte elements: none negative, and neither the element
// count nor the byte count overflows 64 bits (count() and nbytes() are computed from them, and the
// allocation, DLPack export and __cuda_array_interface__ all trust them).
void check_extents(const int64_t* shape, int rank, size_t itemsize)
{
    int64_t elems = 1;
    for (int k = 0; k < rank; ++k) {
        if (shape[k] < 0) fail("negative extent");
        if (__builtin_mul_overflow(elems, shape[k], &elems)) fail("the shape's element count overflows 64 bits");
    }
    int64_t bytes = 0;
    if (__builtin_mul_overflow(elems, (int64_t)itemsize, &bytes)) fail("the shape's byte count overflows 64 bits");
}

void set_dense(ArrayObj* a, const std::vector<int64_t>& shape)
{
    if ((int)shape.size() > kRank)
        fail("rank " + std::to_string(shape.size()) + " exceeds " + std::to_string(kRank));
    check_extents(shape.data(), (int)shape.size(), item(a->dt));
    a->rank = (int)shape.size();
    int64_t acc = 1;
    for (int k = a->rank - 1; k >= 0; --k) {
        if (shape[k] < 0) fail("negative extent");
        a->ext[k] = shape[k];
        a->str[k] = acc;
        acc *= std::max<int64_t>(shape[k], 1);
    }
}

ArrayObj* new_array(StoragePtr st, int dt)
{
    ArrayObj* a = PyObject_New(ArrayObj, &ArrayType);
    if (!a) propagate();
    new (&a->st) StoragePtr(std::move(st));
    a->dt = dt;
    a->rank = 0;
    return a;
}

void array_dealloc(PyObject* self)
{
    auto* a = reinterpret_cast<ArrayObj*>(self);
    a->st.~StoragePtr();
    Py_TYPE(self)->tp_free(self);
}

void* stream_handle(PyObject* s)
{
    if (!s || s == Py_None) return nullptr;
    if (PyLong_Check(s)) {
        const long long h = PyLong_AsLongLong(s);
        if (h == -1 && PyErr_Occurred()) propagate();
        return reinterpret_cast<void*>((intptr_t)h);
    }
    PyObject* v = nullptr;
    if (PyObject_HasAttrString(s, "__cuda_stream__")) { // the stream protocol: (version, handle)
        PyObject* proto = PyObject_GetAttrString(s, "__cuda_stream__");
        if (!proto) propagate();
        PyObject* pair = PyCallable_Check(proto) ? PyObject_CallNoArgs(proto) : (Py_INCREF(proto), proto);
        Py_DECREF(proto);
        if (!pair) propagate();
        v = PySequence_GetItem(pair, 1);
        Py_DECREF(pair);
    } else if (PyObject_HasAttrString(s, "cuda_stream")) { // torch.cuda.Stream
        v = PyObject_GetAttrString(s, "cuda_stream");
    } else {
        fail("stream= wants an int handle, an object with __cuda_stream__, or a torch stream");
    }
    if (!v) propagate();
    const long long h = PyLong_AsLongLong(v);
    Py_DECREF(v);
    if (h == -1 && PyErr_Occurred()) propagate();
    return reinterpret_cast<void*>((intptr_t)h);
}

// fill `bytes` at `ptr` with the `w`-byte element `raw`
void fill_bytes(int kind, int dev, void* ptr, size_t bytes, const unsigned char* raw, size_t w, void* stream)
{
    if (kind == DEV) {
        check(seam()->fill(kind, dev, ptr, bytes, raw, w, stream), "fill");
        return;
    }
    auto* p = static_cast<unsigned char*>(ptr);
    const size_t first = std::min(bytes, (size_t)64);
    for (size_t i = 0; i < first; ++i) p[i] = raw[i % w];
    for (size_t done = first; done < bytes;) { // doubling copies of the filled prefix (64 is a multiple of w)
        const size_t n = std::min(done, bytes - done);
        std::memcpy(p + done, p, n);
        done += n;
    }
}

PyObject* make_array(int kind, int dev, const std::vector<int64_t>& shape, int dt, PyObject* stream_obj, bool zero)
{
    auto st = std::make_shared<Storage>();
    ArrayObj* a = new_array(st, dt);
    try {
        set_dense(a, shape);
        const size_t bytes = nbytes(a);
        st->kind = kind;
        st->dev = dev;
        st->bytes = bytes;
        st->stream = stream_handle(stream_obj);
        if (stream_obj && stream_obj != Py_None) {
            Py_INCREF(stream_obj);
            st->stream_obj = stream_obj;
        }
        if (bytes) {
            if (kind == CPU) {
                st->host = std::aligned_alloc(64, (bytes + 63) / 64 * 64);
                if (!st->host) fail_as(PyExc_MemoryError, "host allocation of " + std::to_string(bytes) + " bytes");
                st->ptr = st->host;
            } else {
                check(seam()->chunk_alloc(kind, dev, bytes, st->stream, &st->chunk, &st->ptr), "alloc");
            }
            if (zero) {
                const unsigned char z = 0;
                fill_bytes(kind, dev, st->ptr, bytes, &z, 1, st->stream);
            }
        }
    } catch (PyErrorSet&) {
        Py_DECREF(a);
        throw;
    }
    return reinterpret_cast<PyObject*>(a);
}

// ------------------------------------------------------------------ Array: properties
PyObject* tuple_of(const int64_t* v, int n, int64_t scale = 1)
{
    PyObject* t = PyTuple_New(n);
    if (!t) propagate();
    for (int k = 0; k < n; ++k) PyTuple_SET_ITEM(t, k, PyLong_FromLongLong(v[k] * scale));
    return t;
}

// Every entry point turns a PyErrorSet (or a C++ exception) into a Python error return.
template<class F>
PyObject* guard(F&& f)
{
    try {
        return f();
    } catch (PyErrorSet&) {
        return nullptr;
    } catch (const std::bad_alloc&) {
        return PyErr_NoMemory();
    } catch (const std::exception& e) {
        PyErr_SetString(g_error, e.what());
        return nullptr;
    }
}

PyObject* device_object(const ArrayObj* a)
{
    if (!g_device_cls) fail("aether._core is not initialised (aether imports it)");
    const int k = a->st->kind;
    return PyObject_CallFunction(g_device_cls, "si", k == DEV ? "cuda" : k == PINNED ? "pinned" : "cpu",
                                 k == DEV ? a->st->dev : 0);
}

PyObject* get_shape(PyObject* s, void*) { return guard([&] { auto* a = (ArrayObj*)s; return tuple_of(a->ext, a->rank); }); }
PyObject* get_strides(PyObject* s, void*)
{
    return guard([&] { auto* a = (ArrayObj*)s; return tuple_of(a->str, a->rank, (int64_t)item(a->dt)); });
}
PyObject* get_dtype(PyObject* s, void*)
{
    return guard([&]() -> PyObject* {
        if (!g_dtype_cls) fail("aether._core is not initialised (aether imports it)");
        return PyObject_CallFunction(g_dtype_cls, "s", kType[((ArrayObj*)s)->dt].name);
    });
}
PyObject* get_ndim(PyObject* s, void*) { return PyLong_FromLong(((ArrayObj*)s)->rank); }
PyObject* get_size(PyObject* s, void*) { return PyLong_FromLongLong(count((ArrayObj*)s)); }
PyObject* get_nbytes(PyObject* s, void*) { return PyLong_FromSize_t(nbytes((ArrayObj*)s)); }
PyObject* get_ptr(PyObject* s, void*) { return PyLong_FromVoidPtr(((ArrayObj*)s)->st->ptr); }
PyObject* get_device(PyObject* s, void*) { return guard([&] { return device_object((ArrayObj*)s); }); }
PyObject* get_contiguous(PyObject* s, void*) { return PyBool_FromLong(contiguous((ArrayObj*)s)); }

PyObject* get_cai(PyObject* s, void*)
{
    return guard([&]() -> PyObject* {
        auto* a = (ArrayObj*)s;
        if (a->st->kind != DEV) fail_as(PyExc_AttributeError, "only device arrays have __cuda_array_interface__");
        if (!kType[a->dt].typestr) fail_as(PyExc_AttributeError, "this dtype has no __cuda_array_interface__ typestr");
        PyObject* shape = tuple_of(a->ext, a->rank);
        PyObject* strides = contiguous(a) ? (Py_INCREF(Py_None), Py_None) : tuple_of(a->str, a->rank, (int64_t)item(a->dt));
        PyObject* stream = a->st->stream ? PyLong_FromVoidPtr(a->st->stream) : (Py_INCREF(Py_None), Py_None);
        PyObject* d = Py_BuildValue("{s:N,s:s,s:(N,O),s:i,s:N,s:N}", "shape", shape, "typestr", kType[a->dt].typestr,
                                    "data", PyLong_FromVoidPtr(a->st->ptr), Py_False, "version", 3, "strides",
                                    strides, "stream", stream);
        return d;
    });
}

// ------------------------------------------------------------------ Array: methods
void do_fill(ArrayObj* a, PyObject* value, PyObject* stream, bool zero)
{
    need_contiguous(a, zero ? "zero" : "fill");
    const size_t n = nbytes(a);
    if (!n) return;
    unsigned char raw[8] = {0};
    if (!zero) pack(a->dt, value, raw);
    const size_t w = zero ? 1 : item(a->dt);
    if (a->st->kind == PINNED && stream && stream != Py_None)
        check(seam()->stream_sync(a->st->dev, stream_handle(stream)), "fill");
    void* h = a->st->kind != DEV ? nullptr : (stream && stream != Py_None) ? stream_handle(stream) : a->st->stream;
    fill_bytes(a->st->kind, a->st->dev, a->st->ptr, n, raw, w, h);
}

PyObject* m_zero(PyObject* self, PyObject* args, PyObject* kw)
{
    return guard([&] {
        static const char* names[] = {"stream", nullptr};
        PyObject* stream = Py_None;
        if (!PyArg_ParseTupleAndKeywords(args, kw, "|O", const_cast<char**>(names), &stream)) propagate();
        do_fill((ArrayObj*)self, nullptr, stream, true);
        Py_INCREF(self);
        return self;
    });
}

PyObject* m_fill(PyObject* self, PyObject* args, PyObject* kw)
{
    return guard([&] {
        static const char* names[] = {"value", "stream", nullptr};
        PyObject *value = nullptr, *stream = Py_None;
        if (!PyArg_ParseTupleAndKeywords(args, kw, "O|O", const_cast<char**>(names), &value, &stream)) propagate();
        do_fill((ArrayObj*)self, value, stream, false);
        Py_INCREF(self);
        return self;
    });
}

PyObject* m_set_flat(PyObject* self, PyObject* args)
{
    return guard([&]() -> PyObject* {
        Py_ssize_t i;
        PyObject* v;
        if (!PyArg_ParseTuple(args, "nO", &i, &v)) propagate();
        auto* a = (ArrayObj*)self;
        if (!host_side(a) || i < 0 || i >= count(a) || !contiguous(a))
            fail("_set_flat: contiguous host arrays only, index in range");
        pack(a->dt, v, static_cast<unsigned char*>(a->st->ptr) + (size_t)i * item(a->dt));
        Py_RETURN_NONE;
    });
}

PyObject* m_synchronize(PyObject* self, PyObject*)
{
    return guard([&]() -> PyObject* {
        auto* a = (ArrayObj*)self;
        if (a->st->kind != CPU) check(seam()->stream_sync(a->st->dev, a->st->stream), "synchronize");
        Py_RETURN_NONE;
    });
}

int32_t dl_device_type(const ArrayObj* a)
{
    return a->st->kind == DEV ? kDLROCM : a->st->kind == PINNED ? kDLROCMHost : kDLCPU;
}

PyObject* m_dlpack_device(PyObject* self, PyObject*)
{
    auto* a = (ArrayObj*)self;
    return Py_BuildValue("(ii)", dl_device_type(a), a->st->kind == DEV ? a->st->dev : 0);
}

// The manager context of an export: it holds the storage (and the shape / strides the tensor points at).
struct Export {
    StoragePtr st;
    int64_t shape[kRank];
    int64_t strides[kRank];
};

template<class M>
void export_deleter(M* m)
{
    auto* e = static_cast<Export*>(m->manager_ctx);
    delete e; // releases the storage (taking the GIL itself for what Python owns)
    delete m;
    --g_exports;
}

template<class M>
void capsule_destructor(PyObject* cap)
{
    // a consumer renames the capsule "used_..." and owns the tensor; otherwise it was never consumed
    const char* name = PyCapsule_GetName(cap);
    if (!name || std::strncmp(name, "used_", 5) == 0) return;
    auto* m = static_cast<M*>(PyCapsule_GetPointer(cap, name));
    if (m && m->deleter) m->deleter(m);
    else PyErr_Clear();
}

template<class M>
M* make_export(const ArrayObj* a)
{
    auto* e = new Export{a->st, {}, {}};
    for (int k = 0; k < a->rank; ++k) {
        e->shape[k] = a->ext[k];
        e->strides[k] = a->str[k];
    }
    auto* m = new M{};
    DLTensor& t = m->dl_tensor;
    t.data = a->st->ptr;
    t.device = {dl_device_type(a), a->st->kind == DEV ? a->st->dev : 0};
    t.ndim = a->rank;
    t.dtype = {kType[a->dt].code, kType[a->dt].bits, 1};
    t.shape = e->shape;
    t.strides = e->strides;
    t.byte_offset = 0;
    m->manager_ctx = e;
    m->deleter = &export_deleter<M>;
    ++g_exports;
    return m;
}

PyObject* m_dlpack(PyObject* self, PyObject* args, PyObject* kw)
{
    return guard([&]() -> PyObject* {
        static const char* names[] = {"stream", "max_version", "dl_device", "copy", nullptr};
        PyObject *stream = Py_None, *max_version = Py_None, *dl_device = Py_None, *copy = Py_None;
        if (!PyArg_ParseTupleAndKeywords(args, kw, "|$OOOO", const_cast<char**>(names), &stream, &max_version,
                                         &dl_device, &copy))
            propagate();
        auto* a = (ArrayObj*)self;
        if (copy != Py_None && PyObject_IsTrue(copy)) fail_as(PyExc_BufferError, "aether arrays do not copy on export (copy=True)");
        if (dl_device != Py_None) {
            int dtype_ = 0, did = 0;
            if (!PyArg_ParseTuple(dl_device, "ii", &dtype_, &did)) propagate();
            if (dtype_ != dl_device_type(a) || did != (a->st->kind == DEV ? a->st->dev : 0))
                fail_as(PyExc_BufferError, "cannot export to another device");
        }
        if (stream != Py_None && a->st->kind == DEV) {
            // ROCm stream codes: 0 the default (null) stream, -1 no synchronisation, 1 and 2 unused,
            // anything else a stream handle
            const long long s = PyLong_AsLongLong(stream);
            if (s == -1 && PyErr_Occurred()) propagate();
            if (s == 1 || s == 2) fail_as(PyExc_ValueError, "__dlpack__(stream=" + std::to_string(s) + ") is not a ROCm stream code");
            if (s != -1) {
                void* consumer = reinterpret_cast<void*>((intptr_t)s);
                if (consumer != a->st->stream)
                    check(seam()->stream_wait(a->st->dev, a->st->stream, consumer), "dlpack stream ordering");
            }
        }
        long major = 0;
        if (max_version != Py_None) {
            PyObject* m0 = PySequence_GetItem(max_version, 0);
            if (!m0) propagate();
            major = PyLong_AsLong(m0);
            Py_DECREF(m0);
            if (major == -1 && PyErr_Occurred()) propagate();
        }
        if (major >= 1) {
            auto* m = make_export<DLManagedTensorVersioned>(a);
            m->version = {1, 0};
            m->flags = 0;
            PyObject* cap = PyCapsule_New(m, "dltensor_versioned", &capsule_destructor<DLManagedTensorVersioned>);
            if (!cap) {
                m->deleter(m);
                propagate();
            }
            return cap;
        }
        auto* m = make_export<DLManagedTensor>(a);
        PyObject* cap = PyCapsule_New(m, "dltensor", &capsule_destructor<DLManagedTensor>);
        if (!cap) {
            m->deleter(m);
            propagate();
        }
        return cap;
    });
}

PyObject* array_repr(PyObject* self)
{
    return guard([&]() -> PyObject* {
        auto* a = (ArrayObj*)self;
        std::string shape = "(";
        for (int k = 0; k < a->rank; ++k) shape += std::to_string(a->ext[k]) + (a->rank == 1 ? "," : k + 1 < a->rank ? ", " : "");
        shape += ")";
        PyObject* dev = device_object(a);
        if (!dev) propagate();
        PyObject* ds = PyObject_Str(dev);
        Py_DECREF(dev);
        if (!ds) propagate();
        PyObject* r = PyUnicode_FromFormat("aether.Array(shape=%s, dtype=%s, device='%U')", shape.c_str(),
                                           kType[a->dt].name, ds);
        Py_DECREF(ds);
        return r;
    });
}

// ------------------------------------------------------------------ buffer protocol (host memory)
int array_getbuffer(PyObject* self, Py_buffer* view, int flags)
{
    auto* a = (ArrayObj*)self;
    view->obj = nullptr;
    if (!host_side(a)) {
        PyErr_SetString(PyExc_BufferError, "device memory has no host buffer");
        return -1;
    }
    if (!kType[a->dt].format) {
        PyErr_SetString(PyExc_BufferError, "this dtype has no buffer format");
        return -1;
    }
    const bool dense = contiguous(a);
    if (!(flags & PyBUF_STRIDES) && !dense) {
        PyErr_SetString(PyExc_BufferError, "a strided array needs a strided buffer request");
        return -1;
    }
    if ((flags & PyBUF_F_CONTIGUOUS) == PyBUF_F_CONTIGUOUS && a->rank > 1 && !dense) {
        PyErr_SetString(PyExc_BufferError, "the array is not Fortran-contiguous");
        return -1;
    }
    const int r = a->rank;
    auto* mem = new (std::nothrow) Py_ssize_t[2 * (r ? r : 1)];
    if (!mem) {
        PyErr_NoMemory();
        return -1;
    }
    const size_t w = item(a->dt);
    for (int k = 0; k < r; ++k) {
        mem[k] = (Py_ssize_t)a->ext[k];
        mem[r + k] = (Py_ssize_t)(a->str[k] * (int64_t)w);
    }
    view->buf = a->st->ptr;
    view->obj = self;
    Py_INCREF(self);
    view->len = (Py_ssize_t)nbytes(a);
    view->readonly = 0;
    view->itemsize = (Py_ssize_t)w;
    view->format = (flags & PyBUF_FORMAT) ? const_cast<char*>(kType[a->dt].format) : nullptr;
    view->ndim = r;
    view->shape = (flags & PyBUF_ND) ? mem : nullptr;
    view->strides = (flags & PyBUF_STRIDES) ? mem + r : nullptr;
    view->suboffsets = nullptr;
    view->internal = mem;
    return 0;
}

void array_releasebuffer(PyObject*, Py_buffer* view) { delete[] static_cast<Py_ssize_t*>(view->internal); }

PyBufferProcs kArrayBuffer = {array_getbuffer, array_releasebuffer};

PyGetSetDef kArrayGetSet[] = {
    {"shape", get_shape, nullptr, "extents", nullptr},
    {"strides", get_strides, nullptr, "strides in bytes", nullptr},
    {"dtype", get_dtype, nullptr, "the element type", nullptr},
    {"ndim", get_ndim, nullptr, nullptr, nullptr},
    {"size", get_size, nullptr, "number of elements", nullptr},
    {"nbytes", get_nbytes, nullptr, nullptr, nullptr},
    {"ptr", get_ptr, nullptr, "address of element 0 (0 for an empty array)", nullptr},
    {"device", get_device, nullptr, "Device(kind, index)", nullptr},
    {"is_contiguous", get_contiguous, nullptr, "dense row-major", nullptr},
    {"__cuda_array_interface__", get_cai, nullptr, "device arrays only", nullptr},
    {nullptr, nullptr, nullptr, nullptr, nullptr},
};

PyMethodDef kArrayMethods[] = {
    {"zero", (PyCFunction)(void (*)(void))m_zero, METH_VARARGS | METH_KEYWORDS,
     "zero(stream=None): all bytes to zero (a kernel on the device, legal inside a capture)"},
    {"fill", (PyCFunction)(void (*)(void))m_fill, METH_VARARGS | METH_KEYWORDS, "fill(value, stream=None)"},
    {"_set_flat", m_set_flat, METH_VARARGS, nullptr},
    {"synchronize", m_synchronize, METH_NOARGS, "wait for the allocation stream"},
    {"__dlpack_device__", m_dlpack_device, METH_NOARGS, nullptr},
    {"__dlpack__", (PyCFunction)(void (*)(void))m_dlpack, METH_VARARGS | METH_KEYWORDS,
     "__dlpack__(*, stream=None, max_version=None, dl_device=None, copy=None)"},
    {nullptr, nullptr, 0, nullptr},
};

// ------------------------------------------------------------------ imports
// Take ownership of a consumed DLPack tensor as a new array.
template<class M>
PyObject* from_managed(M* m, bool versioned)
{
    auto st = std::make_shared<Storage>();
    if (versioned) st->imported_v = reinterpret_cast<DLManagedTensorVersioned*>(m);
    else st->imported = reinterpret_cast<DLManagedTensor*>(m);
    const DLTensor& t = m->dl_tensor;
    const int dt = type_by_dl(t.dtype);
    const int32_t type = t.device.device_type;
    st->kind = type == kDLCPU ? CPU : (type == kDLROCMHost || type == kDLCUDAHost) ? PINNED : type == kDLROCM ? DEV : -1;
    if (st->kind < 0)
        fail("DLPack device type " + std::to_string(type) + " is not supported (cpu 1, rocm 10, rocm host 11)");
    if (t.ndim < 0 || t.ndim > kRank) fail("DLPack rank " + std::to_string(t.ndim) + " exceeds " + std::to_string(kRank));
    check_extents(t.shape, t.ndim, item(dt));
    st->dev = t.device.device_id;
    st->ptr = t.data ? static_cast<char*>(t.data) + t.byte_offset : nullptr;
    ArrayObj* a = new_array(st, dt);
    a->rank = t.ndim;
    int64_t acc = 1;
    for (int k = t.ndim - 1; k >= 0; --k) {
        a->ext[k] = t.shape[k];
        a->str[k] = t.strides ? t.strides[k] : acc;
        acc *= std::max<int64_t>(t.shape[k], 1);
    }
    return reinterpret_cast<PyObject*>(a);
}

PyObject* f_from_dlpack(PyObject*, PyObject* args, PyObject* kw)
{
    return guard([&]() -> PyObject* {
        static const char* names[] = {"x", "stream", nullptr};
        PyObject *x = nullptr, *stream = Py_None;
        if (!PyArg_ParseTupleAndKeywords(args, kw, "O|O", const_cast<char**>(names), &x, &stream)) propagate();
        PyObject* dev = PyObject_CallMethod(x, "__dlpack_device__", nullptr);
        if (!dev) propagate();
        int type = 0, did = 0;
        const bool ok = PyArg_ParseTuple(dev, "ii", &type, &did);
        Py_DECREF(dev);
        if (!ok) propagate();
        // the consumer's stream: the producer orders its work before it (0 names the null stream)
        PyObject* ask = nullptr;
        if (stream != Py_None && type == kDLROCM) {
            ask = PyLong_FromLongLong((long long)(intptr_t)stream_handle(stream));
            if (!ask) propagate();
        }
        auto call = [&](bool versioned) -> PyObject* {
            PyObject* kwargs = PyDict_New();
            if (!kwargs) return nullptr;
            if (versioned) {
                PyObject* v = Py_BuildValue("(ii)", 1, 0);
                PyDict_SetItemString(kwargs, "max_version", v);
                Py_DECREF(v);
            }
            if (ask) PyDict_SetItemString(kwargs, "stream", ask);
            PyObject* fn = PyObject_GetAttrString(x, "__dlpack__");
            PyObject* empty = PyTuple_New(0);
            PyObject* r = fn && empty ? PyObject_Call(fn, empty, kwargs) : nullptr;
            Py_XDECREF(fn);
            Py_XDECREF(empty);
            Py_DECREF(kwargs);
            return r;
        };
        PyObject* cap = call(true);
        if (!cap && PyErr_ExceptionMatches(PyExc_TypeError)) { // a producer without max_version
            PyErr_Clear();
            cap = call(false);
        }
        Py_XDECREF(ask);
        if (!cap) propagate();
        const bool versioned = PyCapsule_IsValid(cap, "dltensor_versioned") != 0;
        if (!versioned && !PyCapsule_IsValid(cap, "dltensor")) {
            Py_DECREF(cap);
            fail("not a DLPack capsule (or one already consumed)");
        }
        void* p = PyCapsule_GetPointer(cap, versioned ? "dltensor_versioned" : "dltensor");
        PyCapsule_SetName(cap, versioned ? "used_dltensor_versioned" : "used_dltensor");
        Py_DECREF(cap);
        // the tensor is ours now: from_managed's storage owns it, also when the import fails
        if (versioned) {
            auto* m = static_cast<DLManagedTensorVersioned*>(p);
            if (m->version.major > 1) {
                if (m->deleter) m->deleter(m);
                fail("DLPack major version " + std::to_string(m->version.major) + " is newer than 1");
            }
            return from_managed(m, true);
        }
        return from_managed(static_cast<DLManagedTensor*>(p), false);
    });
}

PyObject* f_from_buffer(PyObject*, PyObject* obj)
{
    return guard([&]() -> PyObject* {
        auto* view = new Py_buffer();
        if (PyObject_GetBuffer(obj, view, PyBUF_FORMAT | PyBUF_STRIDES) != 0) {
            delete view;
            propagate();
        }
        auto st = std::make_shared<Storage>();
        st->view = view; // released with the storage, also when this import fails below
        st->kind = CPU;
        st->ptr = view->buf;
        st->bytes = (size_t)view->len;
        Py_INCREF(obj);
        st->keep = obj;
        const int dt = type_by_format(view->format ? view->format : "B");
        if (view->ndim > kRank) fail("buffer rank " + std::to_string(view->ndim) + " exceeds " + std::to_string(kRank));
        {
            std::vector<int64_t> shape(view->ndim);
            for (int k = 0; k < view->ndim; ++k) shape[k] = view->shape[k];
            check_extents(shape.data(), view->ndim, item(dt));
        }
        ArrayObj* a = new_array(st, dt);
        a->rank = view->ndim;
        for (int k = 0; k < view->ndim; ++k) {
            a->ext[k] = view->shape[k];
            if (view->strides[k] % view->itemsize) {
                Py_DECREF(a);
                fail("buffer strides are not a whole number of elements");
            }
            a->str[k] = view->strides[k] / view->itemsize;
        }
        return reinterpret_cast<PyObject*>(a);
    });
}

std::vector<int64_t> int_list(PyObject* seq, const char* what)
{
    PyObject* fast = PySequence_Fast(seq, what);
    if (!fast) propagate();
    std::vector<int64_t> v((size_t)PySequence_Fast_GET_SIZE(fast));
    for (size_t k = 0; k < v.size(); ++k) {
        v[k] = PyLong_AsLongLong(PySequence_Fast_GET_ITEM(fast, (Py_ssize_t)k));
        if (v[k] == -1 && PyErr_Occurred()) {
            Py_DECREF(fast);
            propagate();
        }
    }
    Py_DECREF(fast);
    return v;
}

PyObject* f_from_cai(PyObject*, PyObject* obj)
{
    return guard([&]() -> PyObject* {
        PyObject* cai = PyObject_GetAttrString(obj, "__cuda_array_interface__");
        if (!cai) propagate();
        struct Drop {
            PyObject* o;
            ~Drop() { Py_DECREF(o); }
        } drop{cai};
        PyObject* data = PyDict_GetItemString(cai, "data");
        PyObject* typestr = PyDict_GetItemString(cai, "typestr");
        PyObject* shape = PyDict_GetItemString(cai, "shape");
        if (!data || !typestr || !shape) fail("__cuda_array_interface__ lacks data, typestr or shape");
        PyObject* p0 = PySequence_GetItem(data, 0);
        if (!p0) propagate();
        void* ptr = PyLong_AsVoidPtr(p0);
        Py_DECREF(p0);
        if (PyErr_Occurred()) propagate();
        const char* ts = PyUnicode_AsUTF8(typestr);
        if (!ts) propagate();
        int dt = -1;
        for (int t = 0; t < kTypes; ++t)
            if (kType[t].typestr && std::strcmp(ts, kType[t].typestr) == 0) dt = t;
        if (dt < 0) fail(std::string("__cuda_array_interface__ typestr '") + ts + "' is not supported");
        int dev = 0;
        if (ptr) check(seam()->pointer_device(ptr, &dev), "__cuda_array_interface__ pointer");
        auto st = std::make_shared<Storage>();
        st->kind = DEV;
        st->dev = dev;
        st->ptr = ptr;
        Py_INCREF(obj);
        st->keep = obj;
        ArrayObj* a = new_array(st, dt);
        try {
            set_dense(a, int_list(shape, "shape"));
            PyObject* strides = PyDict_GetItemString(cai, "strides");
            if (strides && strides != Py_None) {
                const auto s = int_list(strides, "strides");
                if ((int)s.size() != a->rank) fail("__cuda_array_interface__ strides do not match the shape");
                for (int k = 0; k < a->rank; ++k) {
                    if (s[k] % (int64_t)item(dt)) fail("__cuda_array_interface__ strides are not whole elements");
                    a->str[k] = s[k] / (int64_t)item(dt);
                }
            }
        } catch (PyErrorSet&) {
            Py_DECREF(a);
            throw;
        }
        return reinterpret_cast<PyObject*>(a);
    });
}

// ------------------------------------------------------------------ module functions
PyObject* f_register(PyObject*, PyObject* args)
{
    PyObject *device, *dtype;
    if (!PyArg_ParseTuple(args, "OO", &device, &dtype)) return nullptr;
    Py_INCREF(device);
    Py_INCREF(dtype);
    Py_XSETREF(g_device_cls, device);
    Py_XSETREF(g_dtype_cls, dtype);
    Py_RETURN_NONE;
}

PyObject* f_current_device(PyObject*, PyObject*)
{
    return guard([&] {
        int d = 0;
        check(seam()->current_device(&d), "current device");
        return PyLong_FromLong(d);
    });
}

PyObject* f_runtime_version(PyObject*, PyObject*)
{
    return guard([&] {
        int v = 0;
        check(seam()->runtime_version(&v), "runtime version");
        return PyLong_FromLong(v);
    });
}

PyObject* f_plugin(PyObject*, PyObject*)
{
    return guard([&] {
        seam();
        return PyUnicode_FromString("hip");
    });
}

PyObject* f_new(PyObject*, PyObject* args, PyObject* kw)
{
    return guard([&]() -> PyObject* {
        static const char* names[] = {"kind", "dev", "shape", "dtype", "stream", "zero", nullptr};
        int kind = 0, dev = 0, zero = 0;
        PyObject *shape, *stream = Py_None;
        const char* dtype;
        if (!PyArg_ParseTupleAndKeywords(args, kw, "iiOsOp", const_cast<char**>(names), &kind, &dev, &shape, &dtype,
                                         &stream, &zero))
            propagate();
        if (kind != CPU && kind != DEV && kind != PINNED) fail("unknown memory kind");
        return make_array(kind, dev, int_list(shape, "shape"), type_by_name(dtype), stream, zero != 0);
    });
}

PyObject* f_copy(PyObject*, PyObject* args, PyObject* kw)
{
    return guard([&]() -> PyObject* {
        static const char* names[] = {"src", "dst", "stream", nullptr};
        PyObject *so, *dobj, *stream = Py_None;
        if (!PyArg_ParseTupleAndKeywords(args, kw, "OO|O", const_cast<char**>(names), &so, &dobj, &stream)) propagate();
        ArrayObj* s = as_array(so, "copy: src");
        ArrayObj* d = as_array(dobj, "copy: dst");
        if (s->dt != d->dt || count(s) != count(d))
            fail(std::string("copy: shapes/dtypes differ (") + kType[s->dt].name + " x " + std::to_string(count(s)) +
                 " into " + kType[d->dt].name + " x " + std::to_string(count(d)) + ")");
        if (!contiguous(s) || !contiguous(d)) fail("copy needs contiguous arrays");
        const size_t n = nbytes(s);
        if (!n) Py_RETURN_NONE;
        const bool with_stream = stream != Py_None;
        if (host_side(s) && host_side(d)) {
            if (with_stream && (s->st->kind == PINNED || d->st->kind == PINNED))
                check(seam()->stream_sync(0, stream_handle(stream)), "copy");
            std::memmove(d->st->ptr, s->st->ptr, n);
            Py_RETURN_NONE;
        }
        const AetherMemoryV1* t = seam();
        const int dev = host_side(d) ? s->st->dev : d->st->dev;
        void* h = stream_handle(stream);
        check(t->copy(d->st->kind, d->st->dev, d->st->ptr, s->st->kind, s->st->dev, s->st->ptr, n, h), "copy");
        if (!with_stream) check(t->stream_sync(dev, nullptr), "copy");
        Py_RETURN_NONE;
    });
}

PyObject* f_drain(PyObject*, PyObject*)
{
    if (g_seam) g_seam->drain();
    Py_RETURN_NONE;
}

PyObject* f_exports_live(PyObject*, PyObject*) { return PyLong_FromLong(g_exports.load()); }

// ------------------------------------------------------------------ FFT plans
// A fixed transform of dense interleaved-pair arrays over the seam's plans. A device plan runs on its
// device; a host plan (cpu or pinned arrays) runs the same device plan on the current device, staging
// the data through device memory for each call (the product path has no host transform).
struct PlanObj {
    PyObject_HEAD
    int kind, dt, norm, forward, host, dev;
    std::vector<int64_t>* shape; // complex shape (c2c) or the real side's
    std::vector<int64_t>* half;  // the complex side
    void* plan;
};
PyTypeObject PlanType = {PyVarObject_HEAD_INIT(nullptr, 0)};

std::vector<int64_t> plan_in(const PlanObj* p)
{
    auto s = p->kind == AETHER_FFT_C2R ? *p->half : *p->shape;
    if (p->kind != AETHER_FFT_R2C) s.push_back(2);
    return s;
}
std::vector<int64_t> plan_out(const PlanObj* p)
{
    auto s = p->kind == AETHER_FFT_R2C ? *p->half : *p->shape;
    if (p->kind != AETHER_FFT_C2R) s.push_back(2);
    return s;
}

void plan_dealloc(PyObject* self)
{
    auto* p = reinterpret_cast<PlanObj*>(self);
    if (p->plan && g_seam) g_seam->fft_destroy(p->plan);
    delete p->shape;
    delete p->half;
    Py_TYPE(self)->tp_free(self);
}

PyObject* f_fft_plan(PyObject*, PyObject* args, PyObject* kw)
{
    return guard([&]() -> PyObject* {
        static const char* names[] = {"kind",    "dtype", "device_kind", "device_index", "shape",
                                      "axes",    "forward", "norm",      "stream",       nullptr};
        int kind, dk, di, forward, norm;
        const char* dtype;
        PyObject *shape_o, *axes_o, *stream = Py_None;
        if (!PyArg_ParseTupleAndKeywords(args, kw, "isiiOOpi|O", const_cast<char**>(names), &kind, &dtype, &dk, &di,
                                         &shape_o, &axes_o, &forward, &norm, &stream))
            propagate();
        const int dt = type_by_name(dtype);
        if (dt != F64 && dt != F32) fail(std::string("fft wants float32 or float64, not ") + dtype);
        if (kind < 0 || kind > 2) fail("fft: unknown transform kind");
        if (norm < 0 || norm > 2) fail("fft: unknown norm");
        std::vector<int64_t> shape = int_list(shape_o, "shape");
        std::vector<int64_t> ax64 = int_list(axes_o, "axes");
        if (ax64.empty()) fail("fft: no axes to transform");
        for (size_t i = 0; i < ax64.size(); ++i) {
            if (ax64[i] < 0 || ax64[i] >= (int64_t)shape.size()) fail("fft: axis out of range");
            for (size_t j = 0; j < i; ++j)
                if (ax64[j] == ax64[i]) fail("fft: repeated axis");
        }
        for (int64_t v : shape)
            if (v < 0) fail("fft: negative extent");
        std::vector<int32_t> axes(ax64.begin(), ax64.end());
        const bool host = dk != DEV;
        int dev = di;
        if (host) check(seam()->current_device(&dev), "current device");
        void* h = nullptr;
        const int code = seam()->fft_plan(dev, kind, dt == F64, norm, forward ? 1 : 0, (int)shape.size(), shape.data(),
                                          (int)axes.size(), axes.data(), host ? nullptr : stream_handle(stream), &h);
        if (code == 2) fail_as(g_fft_unavailable, seam()->last_error());
        check(code, "fft plan");
        PlanObj* p = PyObject_New(PlanObj, &PlanType);
        if (!p) {
            seam()->fft_destroy(h);
            propagate();
        }
        p->kind = kind;
        p->dt = dt;
        p->norm = norm;
        p->forward = forward ? 1 : 0;
        p->host = host ? 1 : 0;
        p->dev = dev;
        p->shape = new std::vector<int64_t>(shape);
        p->half = new std::vector<int64_t>(shape);
        if (kind != AETHER_FFT_C2C) (*p->half)[axes.back()] = shape[axes.back()] / 2 + 1;
        p->plan = h;
        return reinterpret_cast<PyObject*>(p);
    });
}

std::string shape_text(const std::vector<int64_t>& s)
{
    std::string w = "(";
    for (size_t i = 0; i < s.size(); ++i) w += (i ? ", " : "") + std::to_string(s[i]);
    return w + (s.size() == 1 ? ",)" : ")");
}

void plan_check(const PlanObj* p, const ArrayObj* a, const std::vector<int64_t>& want, const char* which)
{
    if (a->dt != p->dt)
        fail(std::string("fft: the ") + which + " is " + kType[a->dt].name + "; the plan wants " + kType[p->dt].name);
    const bool on_device = a->st->kind == DEV;
    if (p->host ? on_device : (!on_device || a->st->dev != p->dev))
        fail(std::string("fft: the ") + which + " is not on the plan's device");
    need_contiguous(a, "fft");
    bool same = a->rank == (int)want.size();
    for (int k = 0; same && k < a->rank; ++k) same = a->ext[k] == want[(size_t)k];
    if (!same) fail(std::string("fft: the ") + which + " has the wrong shape; the plan wants " + shape_text(want));
}

PyObject* plan_execute(PyObject* self, PyObject* args, PyObject* kw)
{
    return guard([&]() -> PyObject* {
        static const char* names[] = {"input", "output", "stream", nullptr};
        PyObject *io, *oo, *stream = Py_None;
        if (!PyArg_ParseTupleAndKeywords(args, kw, "OO|O", const_cast<char**>(names), &io, &oo, &stream)) propagate();
        auto* p = reinterpret_cast<PlanObj*>(self);
        ArrayObj* in = as_array(io, "fft: the input");
        ArrayObj* out = as_array(oo, "fft: the output");
        plan_check(p, in, plan_in(p), "input");
        plan_check(p, out, plan_out(p), "output");
        if (p->kind != AETHER_FFT_C2C && in->st->ptr == out->st->ptr && in->st->ptr)
            fail("fft: a real transform needs distinct input and output arrays");
        if (!count(in) || !count(out)) Py_RETURN_NONE;
        const AetherMemoryV1* t = seam();
        if (!p->host) {
            void* h = stream == Py_None ? in->st->stream : stream_handle(stream);
            check(t->fft_exec(p->plan, in->st->ptr, out->st->ptr, h), "fft");
            Py_RETURN_NONE;
        }
        // host arrays: through device staging on the null stream, synchronously
        const size_t ib = nbytes(in), ob = nbytes(out);
        struct Chunks {
            const AetherMemoryV1* t;
            void* c[2] = {nullptr, nullptr};
            ~Chunks()
            {
                for (void* x : c)
                    if (x) t->chunk_release(x);
            }
        } ch{t};
        void *din = nullptr, *dout = nullptr;
        check(t->chunk_alloc(DEV, p->dev, ib, nullptr, &ch.c[0], &din), "fft staging");
        check(t->chunk_alloc(DEV, p->dev, ob, nullptr, &ch.c[1], &dout), "fft staging");
        check(t->copy(DEV, p->dev, din, in->st->kind, 0, in->st->ptr, ib, nullptr), "fft staging");
        check(t->fft_exec(p->plan, din, dout, nullptr), "fft");
        check(t->copy(out->st->kind, 0, out->st->ptr, DEV, p->dev, dout, ob, nullptr), "fft staging");
        check(t->stream_sync(p->dev, nullptr), "fft");
        Py_RETURN_NONE;
    });
}

PyObject* plan_get_kind(PyObject* s, void*) { return PyLong_FromLong(((PlanObj*)s)->kind); }
PyObject* plan_get_in(PyObject* s, void*)
{
    return guard([&] { auto v = plan_in((PlanObj*)s); return tuple_of(v.data(), (int)v.size()); });
}
PyObject* plan_get_out(PyObject* s, void*)
{
    return guard([&] { auto v = plan_out((PlanObj*)s); return tuple_of(v.data(), (int)v.size()); });
}

PyGetSetDef kPlanGetSet[] = {
    {"kind", plan_get_kind, nullptr, "0 c2c, 1 r2c, 2 c2r", nullptr},
    {"input_shape", plan_get_in, nullptr, nullptr, nullptr},
    {"output_shape", plan_get_out, nullptr, nullptr, nullptr},
    {nullptr, nullptr, nullptr, nullptr, nullptr},
};
PyMethodDef kPlanMethods[] = {
    {"execute", (PyCFunction)(void (*)(void))plan_execute, METH_VARARGS | METH_KEYWORDS,
     "execute(input, output, stream=None): a device plan allocates nothing and may run inside a capture"},
    {nullptr, nullptr, 0, nullptr},
};

PyMethodDef kModuleMethods[] = {
    {"_register", f_register, METH_VARARGS, nullptr},
    {"_current_device", f_current_device, METH_NOARGS, nullptr},
    {"cuda_runtime_version", f_runtime_version, METH_NOARGS, "the HIP runtime version"},
    {"cuda_plugin", f_plugin, METH_NOARGS, "load the device library; its runtime's name"},
    {"_new", (PyCFunction)(void (*)(void))f_new, METH_VARARGS | METH_KEYWORDS, nullptr},
    {"_from_dlpack", (PyCFunction)(void (*)(void))f_from_dlpack, METH_VARARGS | METH_KEYWORDS, nullptr},
    {"_from_buffer", f_from_buffer, METH_O, nullptr},
    {"_from_cai", f_from_cai, METH_O, nullptr},
    {"_copy", (PyCFunction)(void (*)(void))f_copy, METH_VARARGS | METH_KEYWORDS, nullptr},
    {"_drain", f_drain, METH_NOARGS, nullptr},
    {"_exports_live", f_exports_live, METH_NOARGS, nullptr},
    {"_fft_plan", (PyCFunction)(void (*)(void))f_fft_plan, METH_VARARGS | METH_KEYWORDS, nullptr},
    {nullptr, nullptr, 0, nullptr},
};

PyModuleDef kModule = {PyModuleDef_HEAD_INIT, "_core", "aether arrays over the HIP seam", -1, kModuleMethods,
                       nullptr, nullptr, nullptr, nullptr};

} // namespace

PyMODINIT_FUNC PyInit__core(void)
{
    ArrayType.tp_name = "aether.Array";
    ArrayType.tp_basicsize = sizeof(ArrayObj);
    ArrayType.tp_dealloc = array_dealloc;
    ArrayType.tp_repr = array_repr;
    ArrayType.tp_as_buffer = &kArrayBuffer;
    ArrayType.tp_flags = Py_TPFLAGS_DEFAULT;
    ArrayType.tp_doc = "One buffer on one device (cpu, pinned or cuda); made by aether.empty/zeros/full/asarray/from_dlpack.";
    ArrayType.tp_methods = kArrayMethods;
    ArrayType.tp_getset = kArrayGetSet;
    PlanType.tp_name = "aether.FftPlan";
    PlanType.tp_basicsize = sizeof(PlanObj);
    PlanType.tp_dealloc = plan_dealloc;
    PlanType.tp_flags = Py_TPFLAGS_DEFAULT;
    PlanType.tp_methods = kPlanMethods;
    PlanType.tp_getset = kPlanGetSet;
    if (PyType_Ready(&ArrayType) < 0 || PyType_Ready(&PlanType) < 0) return nullptr;
    PyObject* m = PyModule_Create(&kModule);
    if (!m) return nullptr;
    g_error = PyErr_NewException("aether.Error", PyExc_RuntimeError, nullptr);
    g_unavailable = PyErr_NewException("aether.CudaUnavailable", g_error, nullptr);
    g_fft_unavailable = PyErr_NewException("aether.FftUnavailable", g_error, nullptr);
    Py_INCREF(&ArrayType);
    Py_INCREF(&PlanType);
    if (!g_error || !g_unavailable || !g_fft_unavailable || PyModule_AddObject(m, "Error", (Py_INCREF(g_error), g_error)) ||
        PyModule_AddObject(m, "CudaUnavailable", (Py_INCREF(g_unavailable), g_unavailable)) ||
        PyModule_AddObject(m, "FftUnavailable", (Py_INCREF(g_fft_unavailable), g_fft_unavailable)) ||
        PyModule_AddObject(m, "Array", reinterpret_cast<PyObject*>(&ArrayType)) ||
        PyModule_AddObject(m, "FftPlan", reinterpret_cast<PyObject*>(&PlanType))) {
        Py_DECREF(m);
        return nullptr;
    }
    return m;
}
