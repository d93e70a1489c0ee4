// Checkpoint files: MessagePack with the reference's positional layout.
//
// The reference serialises its whole Simulation with rmp_serde::Serializer (default
// configuration: structs as positional arrays) (simulation.rs:130-176, 448-479;
// fields.rs:55-64, 143-149; opencl.rs:183-192):
//   Simulation = [cells, fluid, tracers, gravity, iteration, couplings]
//   cells      = [Data(flags i32), counts Array1<usize>]
//   fluid      = [Data(f), Data(rho), Data(vel), nil | Data(acc), model, omega]
//   model      = [ws Array1<f32>, qs Array2<i32>]          (velocities.rs:16-19)
//   tracers    = map name -> [Data(g), Data(C), model, omega]
//   gravity    = nil | Data(g)
//   couplings  = [[tracer, alpha f32, c0 f32, Data(gravity)], ...]
//   Data(x)    = [x]            (the device buffer is not serialised)
//   Array      = [1, [dims...], [elements...]]                (ndarray's serde form)
//   omega      = {"BGK": w} | {"TRT": [w+, w-]}               (externally tagged enum)
// The acceleration buffer's host copy is never read back in the reference, so it is
// written as zeros.  The ndarray and enum encodings follow those crates' documented serde
// forms; no checkpoint written by the reference is available to compare bytes against.
#pragma once

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include "host_init.hpp"

namespace bolt_ckpt {

struct TracerRef {
    std::string name;
    int q;
    int omega_kind;
    float omega_a, omega_b;
    const std::vector<float>* g;
    const std::vector<float>* C;
};

struct CouplingRef {
    std::string tracer;
    float alpha, c0;
    std::vector<float> gravity;
};

struct State {
    std::vector<int64_t> counts;
    const std::vector<int32_t>* flags = nullptr;
    const std::vector<float>* f = nullptr;
    const std::vector<float>* rho = nullptr;
    const std::vector<float>* vel = nullptr;
    int q = 0;
    int omega_kind = 0;
    float omega_a = 0, omega_b = 0;
    bool has_acc = false;
    std::vector<TracerRef> tracers;
    bool has_gravity = false;
    std::vector<float> gravity;
    std::vector<CouplingRef> couplings;
    uint64_t iteration = 0;
};

struct LoadedTracer {
    std::string name;
    int q = 0;
    int omega_kind = 0;
    float omega_a = 0, omega_b = 0;
    std::vector<float> g, C;
};

struct Loaded {
    std::vector<int64_t> counts;
    std::vector<int32_t> flags;
    std::vector<float> f, rho, vel;
    int q = 0;
    int omega_kind = 0;
    float omega_a = 0, omega_b = 0;
    std::vector<LoadedTracer> tracers;
    bool has_gravity = false;
    std::vector<float> gravity;
    std::vector<CouplingRef> couplings;
    uint64_t iteration = 0;
};

// ---- writer -------------------------------------------------------------------------
class Writer {
   public:
    explicit Writer(FILE* fp) : fp_(fp) { buf_.reserve(1 << 20); }
    ~Writer() { flush(); }
    void flush() {
        if (!buf_.empty()) fwrite(buf_.data(), 1, buf_.size(), fp_);
        buf_.clear();
    }
    void byte(uint8_t b) {
        buf_.push_back(b);
        if (buf_.size() >= (1u << 20)) flush();
    }
    void be(uint64_t v, int n) {
        for (int i = n - 1; i >= 0; --i) byte((uint8_t)(v >> (8 * i)));
    }
    void nil() { byte(0xc0); }
    void array(uint32_t n) {
        if (n < 16) byte(0x90 | n);
        else if (n < 65536) { byte(0xdc); be(n, 2); }
        else { byte(0xdd); be(n, 4); }
    }
    void map(uint32_t n) {
        if (n < 16) byte(0x80 | n);
        else if (n < 65536) { byte(0xde); be(n, 2); }
        else { byte(0xdf); be(n, 4); }
    }
    void uint(uint64_t v) {
        if (v < 128) byte((uint8_t)v);
        else if (v < 256) { byte(0xcc); be(v, 1); }
        else if (v < 65536) { byte(0xcd); be(v, 2); }
        else if (v < (1ull << 32)) { byte(0xce); be(v, 4); }
        else { byte(0xcf); be(v, 8); }
    }
    void sint(int64_t v) {
        if (v >= 0) { uint((uint64_t)v); return; }
        if (v >= -32) byte((uint8_t)(int8_t)v);
        else if (v >= -128) { byte(0xd0); be((uint8_t)(int8_t)v, 1); }
        else if (v >= -32768) { byte(0xd1); be((uint16_t)(int16_t)v, 2); }
        else if (v >= -2147483648ll) { byte(0xd2); be((uint32_t)(int32_t)v, 4); }
        else { byte(0xd3); be((uint64_t)v, 8); }
    }
    void f32(float x) {
        uint32_t u;
        memcpy(&u, &x, 4);
        byte(0xca);
        be(u, 4);
    }
    void str(const std::string& s) {
        const size_t n = s.size();
        if (n < 32) byte(0xa0 | (uint8_t)n);
        else if (n < 256) { byte(0xd9); be(n, 1); }
        else { byte(0xda); be(n, 2); }
        for (char c : s) byte((uint8_t)c);
    }
    // ndarray: [1, dims, data]
    template <typename F>
    void ndarray(const std::vector<int64_t>& dims, size_t n, F&& elem) {
        array(3);
        uint(1);
        array((uint32_t)dims.size());
        for (auto d : dims) uint((uint64_t)d);
        array((uint32_t)n);
        for (size_t i = 0; i < n; ++i) elem(i);
    }

   private:
    FILE* fp_;
    std::vector<uint8_t> buf_;
};

inline void write_model(Writer& w, int d, int q) {
    const bolt_host::Table* t = bolt_host::table(d, q);
    w.array(2);
    w.ndarray({q}, (size_t)q, [&](size_t i) { w.f32(bolt_host::weight(t, (int)i)); });
    w.ndarray({q, d}, (size_t)q * d, [&](size_t k) { w.sint(t->c[k / d][k % d]); });
}

inline void write_omega(Writer& w, int kind, float a, float b) {
    w.map(1);
    if (kind == 0) {
        w.str("BGK");
        w.f32(a);
    } else {
        w.str("TRT");
        w.array(2);
        w.f32(a);
        w.f32(b);
    }
}

inline void data_f32(Writer& w, const std::vector<int64_t>& dims, const std::vector<float>& v) {
    w.array(1);
    w.ndarray(dims, v.size(), [&](size_t i) { w.f32(v[i]); });
}

inline bool write(const std::string& path, const State& s, std::string* err) {
    FILE* fp = fopen(path.c_str(), "wb");
    if (!fp) {
        *err = "cannot open " + path;
        return false;
    }
    const int d = (int)s.counts.size();
    std::vector<int64_t> cd = s.counts, cq = s.counts, cdd = s.counts, cacc;
    cq.push_back(s.q);
    cdd.push_back(d);
    size_t ncell = 1;
    for (auto c : s.counts) ncell *= (size_t)c;
    {
        Writer w(fp);
        w.array(6);
        // cells
        w.array(2);
        w.array(1);
        w.ndarray(cd, s.flags->size(), [&](size_t i) { w.sint((*s.flags)[i]); });
        w.ndarray({d}, (size_t)d, [&](size_t i) { w.uint((uint64_t)s.counts[i]); });
        // fluid
        w.array(6);
        data_f32(w, cq, *s.f);
        data_f32(w, cd, *s.rho);
        data_f32(w, cdd, *s.vel);
        if (s.has_acc) {
            w.array(1);
            w.ndarray({(int64_t)ncell, d}, ncell * d, [&](size_t) { w.f32(0.0f); });
        } else {
            w.nil();
        }
        write_model(w, d, s.q);
        write_omega(w, s.omega_kind, s.omega_a, s.omega_b);
        // tracers
        w.map((uint32_t)s.tracers.size());
        for (auto& t : s.tracers) {
            w.str(t.name);
            w.array(4);
            std::vector<int64_t> tq = s.counts;
            tq.push_back(t.q);
            data_f32(w, tq, *t.g);
            data_f32(w, cd, *t.C);
            write_model(w, d, t.q);
            write_omega(w, t.omega_kind, t.omega_a, t.omega_b);
        }
        // gravity
        if (s.has_gravity) data_f32(w, {d}, s.gravity);
        else w.nil();
        w.uint(s.iteration);
        // couplings
        w.array((uint32_t)s.couplings.size());
        for (auto& c : s.couplings) {
            w.array(4);
            w.str(c.tracer);
            w.f32(c.alpha);
            w.f32(c.c0);
            data_f32(w, {d}, c.gravity);
        }
    }
    const bool ok = fclose(fp) == 0;
    if (!ok) *err = "write failed for " + path;
    return ok;
}

// ---- reader -------------------------------------------------------------------------
class Reader {
   public:
    explicit Reader(std::vector<uint8_t> b) : b_(std::move(b)) {}
    bool ok() const { return ok_; }
    const std::string& error() const { return err_; }

    uint8_t peek() { return pos_ < b_.size() ? b_[pos_] : (fail("truncated"), 0); }
    uint8_t next() { return pos_ < b_.size() ? b_[pos_++] : (fail("truncated"), 0); }
    uint64_t be(int n) {
        uint64_t v = 0;
        for (int i = 0; i < n; ++i) v = (v << 8) | next();
        return v;
    }
    void fail(const std::string& m) {
        if (ok_) err_ = m + " at byte " + std::to_string(pos_);
        ok_ = false;
        pos_ = b_.size();
    }
    bool is_nil() { return peek() == 0xc0; }
    void nil() {
        if (next() != 0xc0) fail("expected nil");
    }
    uint32_t array() {
        const uint8_t t = next();
        if ((t & 0xf0) == 0x90) return t & 0x0f;
        if (t == 0xdc) return (uint32_t)be(2);
        if (t == 0xdd) return (uint32_t)be(4);
        fail("expected array");
        return 0;
    }
    uint32_t map() {
        const uint8_t t = next();
        if ((t & 0xf0) == 0x80) return t & 0x0f;
        if (t == 0xde) return (uint32_t)be(2);
        if (t == 0xdf) return (uint32_t)be(4);
        fail("expected map");
        return 0;
    }
    int64_t integer() {
        const uint8_t t = next();
        if (t < 0x80) return t;
        if (t >= 0xe0) return (int8_t)t;
        switch (t) {
            case 0xcc: return (int64_t)be(1);
            case 0xcd: return (int64_t)be(2);
            case 0xce: return (int64_t)be(4);
            case 0xcf: return (int64_t)be(8);
            case 0xd0: return (int8_t)be(1);
            case 0xd1: return (int16_t)be(2);
            case 0xd2: return (int32_t)be(4);
            case 0xd3: return (int64_t)be(8);
        }
        fail("expected integer");
        return 0;
    }
    float f32() {
        const uint8_t t = next();
        if (t == 0xca) {
            uint32_t u = (uint32_t)be(4);
            float x;
            memcpy(&x, &u, 4);
            return x;
        }
        if (t == 0xcb) {
            uint64_t u = be(8);
            double x;
            memcpy(&x, &u, 8);
            return (float)x;
        }
        --pos_;
        return (float)integer();
    }
    std::string str() {
        const uint8_t t = next();
        size_t n;
        if ((t & 0xe0) == 0xa0) n = t & 0x1f;
        else if (t == 0xd9) n = be(1);
        else if (t == 0xda) n = be(2);
        else if (t == 0xdb) n = be(4);
        else {
            fail("expected string");
            return {};
        }
        if (pos_ + n > b_.size()) {
            fail("truncated string");
            return {};
        }
        std::string s((const char*)&b_[pos_], n);
        pos_ += n;
        return s;
    }
    // ndarray [1, dims, data]
    template <typename T, typename F>
    std::vector<T> ndarray(std::vector<int64_t>* dims, F&& elem) {
        std::vector<T> out;
        if (array() != 3) fail("ndarray: expected 3 fields");
        if (integer() != 1) fail("ndarray: unknown version");
        const uint32_t nd = array();
        size_t n = 1;
        dims->clear();
        for (uint32_t i = 0; i < nd && ok_; ++i) {
            dims->push_back(integer());
            n *= (size_t)dims->back();
        }
        const uint32_t m = array();
        if (m != n) fail("ndarray: element count does not match shape");
        if (!ok_) return out;
        out.reserve(n);
        for (size_t i = 0; i < n && ok_; ++i) out.push_back(elem());
        return out;
    }
    std::vector<float> data_f32(std::vector<int64_t>* dims) {
        if (array() != 1) fail("Data: expected 1 field");
        return ndarray<float>(dims, [&]() { return f32(); });
    }
    void omega(int* kind, float* a, float* b) {
        if (map() != 1) fail("Omega: expected one variant");
        const std::string v = str();
        if (v == "BGK") {
            *kind = 0;
            *a = *b = f32();
        } else if (v == "TRT") {
            *kind = 1;
            if (array() != 2) fail("Omega::TRT: expected 2 values");
            *a = f32();
            *b = f32();
        } else {
            fail("Omega: unknown variant " + v);
        }
    }
    void model(int* q) {
        if (array() != 2) fail("model: expected [ws, qs]");
        std::vector<int64_t> dims;
        ndarray<float>(&dims, [&]() { return f32(); });
        *q = dims.empty() ? 0 : (int)dims[0];
        ndarray<int64_t>(&dims, [&]() { return integer(); });
    }

   private:
    std::vector<uint8_t> b_;
    size_t pos_ = 0;
    bool ok_ = true;
    std::string err_;
};

inline bool read(const std::string& path, Loaded* out, std::string* err) {
    FILE* fp = fopen(path.c_str(), "rb");
    if (!fp) {
        *err = "cannot open " + path;
        return false;
    }
    std::vector<uint8_t> bytes;
    uint8_t buf[1 << 16];
    size_t n;
    while ((n = fread(buf, 1, sizeof buf, fp)) > 0) bytes.insert(bytes.end(), buf, buf + n);
    fclose(fp);
    Reader r(std::move(bytes));
    std::vector<int64_t> dims;
    if (r.array() != 6) r.fail("Simulation: expected 6 fields");
    // cells
    if (r.array() != 2) r.fail("Cells: expected 2 fields");
    if (r.array() != 1) r.fail("Data: expected 1 field");
    auto flags64 = r.ndarray<int64_t>(&dims, [&]() { return r.integer(); });
    out->flags.assign(flags64.begin(), flags64.end());
    auto counts = r.ndarray<int64_t>(&dims, [&]() { return r.integer(); });
    out->counts = counts;
    // fluid
    if (r.array() != 6) r.fail("Fluid: expected 6 fields");
    out->f = r.data_f32(&dims);
    out->rho = r.data_f32(&dims);
    out->vel = r.data_f32(&dims);
    if (r.is_nil()) r.nil();
    else r.data_f32(&dims);
    r.model(&out->q);
    r.omega(&out->omega_kind, &out->omega_a, &out->omega_b);
    // tracers
    const uint32_t nt = r.map();
    for (uint32_t i = 0; i < nt && r.ok(); ++i) {
        LoadedTracer t;
        t.name = r.str();
        if (r.array() != 4) r.fail("Scalar: expected 4 fields");
        t.g = r.data_f32(&dims);
        t.C = r.data_f32(&dims);
        r.model(&t.q);
        r.omega(&t.omega_kind, &t.omega_a, &t.omega_b);
        out->tracers.push_back(std::move(t));
    }
    if (r.is_nil()) {
        r.nil();
    } else {
        out->has_gravity = true;
        out->gravity = r.data_f32(&dims);
    }
    out->iteration = (uint64_t)r.integer();
    const uint32_t nc = r.array();
    for (uint32_t i = 0; i < nc && r.ok(); ++i) {
        if (r.array() != 4) r.fail("BoussinesqCoupling: expected 4 fields");
        CouplingRef c;
        c.tracer = r.str();
        c.alpha = r.f32();
        c.c0 = r.f32();
        c.gravity = r.data_f32(&dims);
        out->couplings.push_back(std::move(c));
    }
    if (!r.ok()) {
        *err = r.error();
        return false;
    }
    size_t ncell = 1;
    for (auto c : out->counts) ncell *= (size_t)c;
    const size_t d = out->counts.size();
    if (out->flags.size() != ncell || out->rho.size() != ncell || out->vel.size() != ncell * d ||
        out->f.size() != ncell * (size_t)out->q) {
        *err = "inconsistent array sizes";
        return false;
    }
    return true;
}

}  // namespace bolt_ckpt
