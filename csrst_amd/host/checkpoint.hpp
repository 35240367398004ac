// Restart records of the driver: what SaveH5Part::write stores per step
// (SaveH5Part.cpp:79-259) and LoadH5Part restores (LoadH5Part.cpp:60-321),
// in a flat little-endian file that needs no library.  The csrst_h5part tool
// converts it to and from the H5Part (H5hut) layout.
//
//   file   = "CSRSTRS1" record*
//   record = "STEP"
//            int64 Nx, Ny, n, step
//            double time [s], hx, hy, ox, oy          (E mesh spacing, origin)
//            double Ex[Ny][Nx], Ey, Hzx, Hzy, Jx, Jy  (domain cells, i fastest)
//            double x[n], y[n], oldx[n], oldy[n], px[n], py[n], q[n]
//            int64 id[n]
//            int64 nattr, then per attribute: int32 length, name,
//            uint8 kind ('s' string, 'd' double, 'i' integer), value
//            ('s': int32 length + bytes; 'd': double; 'i': int64)
// The attributes are the command-line arguments, as SaveH5Part writes them
// as file attributes (SaveH5Part.cpp:272-352).
#pragma once

#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace csrst {
namespace ckpt {

struct Attr {
    std::string name;
    char kind = 'i';   // 's', 'd', 'i'
    std::string s;
    double d = 0;
    int64_t i = 0;
};

struct Record {
    int64_t Nx = 0, Ny = 0, step = 0;
    double time = 0, hx = 0, hy = 0, ox = 0, oy = 0;
    std::vector<double> E, H, J;   // 2 * Nx * Ny each: x (or zx) plane, then y (or zy)
    std::vector<double> x, y, oldx, oldy, px, py, q;
    std::vector<int64_t> id;
    std::vector<Attr> attrs;
    size_t n() const { return x.size(); }
    const Attr *attr(const std::string &name) const
    {
        for (const Attr &a : attrs)
            if (a.name == name) return &a;
        return nullptr;
    }
};

constexpr char kMagic[9] = "CSRSTRS1";

class Writer {
public:
    // append: keep the records already in path (a new file gets the magic)
    explicit Writer(const std::string &path, bool append)
    {
        bool fresh = true;
        if (append) {
            std::ifstream in(path, std::ios::binary);
            char m[8];
            fresh = !(in && in.read(m, 8) && std::memcmp(m, kMagic, 8) == 0);
        }
        f_.open(path, std::ios::binary | (fresh ? std::ios::trunc : std::ios::app));
        if (!f_) throw std::runtime_error("cannot write " + path);
        if (fresh) f_.write(kMagic, 8);
    }
    void write(const Record &r)
    {
        const size_t cells = (size_t)(r.Nx * r.Ny), n = r.n();
        if (r.E.size() != 2 * cells || r.H.size() != 2 * cells || r.J.size() != 2 * cells)
            throw std::invalid_argument("restart record: field sizes");
        for (const std::vector<double> *v : {&r.y, &r.oldx, &r.oldy, &r.px, &r.py, &r.q})
            if (v->size() != n) throw std::invalid_argument("restart record: particle array sizes");
        if (r.id.size() != n) throw std::invalid_argument("restart record: id size");
        f_.write("STEP", 4);
        i64(r.Nx); i64(r.Ny); i64((int64_t)n); i64(r.step);
        for (double v : {r.time, r.hx, r.hy, r.ox, r.oy}) f64(v);
        for (const std::vector<double> *v : {&r.E, &r.H, &r.J, &r.x, &r.y, &r.oldx, &r.oldy, &r.px, &r.py, &r.q})
            raw(v->data(), v->size() * sizeof(double));
        raw(r.id.data(), n * sizeof(int64_t));
        i64((int64_t)r.attrs.size());
        for (const Attr &a : r.attrs) {
            str(a.name);
            f_.put(a.kind);
            if (a.kind == 's') str(a.s);
            else if (a.kind == 'd') f64(a.d);
            else i64(a.i);
        }
        f_.flush();
        if (!f_) throw std::runtime_error("restart record write failed");
    }

private:
    void raw(const void *p, size_t b) { f_.write(static_cast<const char *>(p), (std::streamsize)b); }
    void i64(int64_t v) { raw(&v, 8); }
    void f64(double v) { raw(&v, 8); }
    void str(const std::string &s)
    {
        const int32_t l = (int32_t)s.size();
        raw(&l, 4);
        raw(s.data(), s.size());
    }
    std::ofstream f_;
};

// All records of a file, in order; throws on a bad magic or a short record.
inline std::vector<Record> read_all(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open restart file " + path);
    char m[8];
    if (!f.read(m, 8) || std::memcmp(m, kMagic, 8) != 0) throw std::runtime_error(path + ": not a restart file");
    auto raw = [&](void *p, size_t b) {
        if (!f.read(static_cast<char *>(p), (std::streamsize)b)) throw std::runtime_error(path + ": truncated");
    };
    auto i64 = [&] { int64_t v; raw(&v, 8); return v; };
    auto f64 = [&] { double v; raw(&v, 8); return v; };
    auto str = [&] {
        int32_t l;
        raw(&l, 4);
        if (l < 0 || l > (1 << 20)) throw std::runtime_error(path + ": bad string length");
        std::string s((size_t)l, '\0');
        raw(&s[0], (size_t)l);
        return s;
    };
    std::vector<Record> out;
    char tag[4];
    while (f.read(tag, 4)) {
        if (std::memcmp(tag, "STEP", 4) != 0) throw std::runtime_error(path + ": bad record tag");
        Record r;
        r.Nx = i64(); r.Ny = i64();
        const int64_t n = i64();
        r.step = i64();
        if (r.Nx <= 0 || r.Ny <= 0 || n < 0 || r.Nx * r.Ny > (int64_t)1 << 34 || n > (int64_t)1 << 34)
            throw std::runtime_error(path + ": bad record sizes");
        r.time = f64(); r.hx = f64(); r.hy = f64(); r.ox = f64(); r.oy = f64();
        const size_t cells = (size_t)(r.Nx * r.Ny);
        for (std::vector<double> *v : {&r.E, &r.H, &r.J}) {
            v->resize(2 * cells);
            raw(v->data(), 2 * cells * sizeof(double));
        }
        for (std::vector<double> *v : {&r.x, &r.y, &r.oldx, &r.oldy, &r.px, &r.py, &r.q}) {
            v->resize((size_t)n);
            raw(v->data(), (size_t)n * sizeof(double));
        }
        r.id.resize((size_t)n);
        raw(r.id.data(), (size_t)n * sizeof(int64_t));
        const int64_t na = i64();
        if (na < 0 || na > 4096) throw std::runtime_error(path + ": bad attribute count");
        for (int64_t k = 0; k < na; ++k) {
            Attr a;
            a.name = str();
            a.kind = (char)f.get();
            if (a.kind == 's') a.s = str();
            else if (a.kind == 'd') a.d = f64();
            else if (a.kind == 'i') a.i = i64();
            else throw std::runtime_error(path + ": bad attribute kind");
            r.attrs.push_back(std::move(a));
        }
        out.push_back(std::move(r));
    }
    if (out.empty()) throw std::runtime_error(path + ": no records");
    return out;
}

}  // namespace ckpt
}  // namespace csrst
