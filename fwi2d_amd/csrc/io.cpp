// On-disk formats of the reference's programs, on the host: SEPlib headers with their
// binaries (sepRead / sepWrite, src/include/IO.hpp:14-176), the plain binary descriptor
// (binRead / binWrite, IO.hpp:180-430), the source-receiver geometry file (readCoord,
// src/cpp/param.cpp:4-77) and the regular acquisition of the shot parameters
// (readParameters, param.cpp:221-244).  Data are returned in single precision, the
// reference's data_t, converted from the file's element size.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "fwi2d_hip.h"

namespace fwi2d {
extern thread_local std::string g_last_error;
}

namespace {

template <class F>
int io_guard(F&& f)
{
    try {
        f();
        fwi2d::g_last_error.clear();
        return 0;
    } catch (const std::exception& ex) {
        fwi2d::g_last_error = ex.what();
        return 1;
    }
}

std::string strip_quotes(std::string v)
{
    if (v.size() >= 2 && (v.front() == '"' || v.front() == '\'') && v.back() == v.front())
        v = v.substr(1, v.size() - 2);
    return v;
}

// key=value pairs of a header, in file order; a later definition of a key replaces an
// earlier one, as the history of a SEPlib header is read.  Values may be quoted (and
// then hold blanks); words without '=' are prose ("sets next:") and are skipped.  With
// skip_hash, lines starting with '#' are comments (binRead, IO.hpp:193-195).
std::map<std::string, std::string> parse_pairs(const std::string& path, bool skip_hash)
{
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error("The input file " + path + " could not be opened");
    std::map<std::string, std::string> kv;
    std::string line;
    while (std::getline(in, line)) {
        if (skip_hash && !line.empty() && line[0] == '#') continue;
        size_t i = 0;
        const size_t n = line.size();
        while (i < n) {
            while (i < n && std::isspace((unsigned char)line[i])) i++;
            size_t k0 = i;
            while (i < n && !std::isspace((unsigned char)line[i]) && line[i] != '=') i++;
            if (i >= n || line[i] != '=') {  // a word without '=': skip it
                while (i < n && !std::isspace((unsigned char)line[i])) i++;
                continue;
            }
            const std::string key = line.substr(k0, i - k0);
            i++;  // '='
            std::string val;
            if (i < n && (line[i] == '"' || line[i] == '\'')) {
                const char q = line[i];
                const size_t e = line.find(q, i + 1);
                if (e == std::string::npos) {
                    val = line.substr(i + 1);
                    i = n;
                } else {
                    val = line.substr(i + 1, e - i - 1);
                    i = e + 1;
                }
            } else {
                size_t v0 = i;
                while (i < n && !std::isspace((unsigned char)line[i])) i++;
                val = line.substr(v0, i - v0);
            }
            if (!key.empty()) kv[key] = val;
        }
    }
    return kv;
}

long long to_ll(const std::string& key, const std::string& v)
{
    char* end = nullptr;
    errno = 0;
    const long long x = std::strtoll(v.c_str(), &end, 10);
    if (end == v.c_str() || errno) throw std::runtime_error("header value " + key + "=" + v + " is not an integer");
    return x;
}

double to_d(const std::string& key, const std::string& v)
{
    char* end = nullptr;
    errno = 0;
    const double x = std::strtod(v.c_str(), &end);
    if (end == v.c_str() || errno) throw std::runtime_error("header value " + key + "=" + v + " is not a number");
    return x;
}

std::string dirname_of(const std::string& p)
{
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? std::string(".") : (s == 0 ? std::string("/") : p.substr(0, s));
}

std::string basename_of(const std::string& p)
{
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? p : p.substr(s + 1);
}

bool file_exists(const std::string& p)
{
    FILE* f = std::fopen(p.c_str(), "rb");
    if (!f) return false;
    std::fclose(f);
    return true;
}

void copy_str(char* dst, size_t cap, const std::string& s)
{
    const size_t m = std::min(cap - 1, s.size());
    std::memcpy(dst, s.data(), m);
    dst[m] = 0;
}

// The hypercube of a header: n1..n6, o, d, labels (defaults 1, 0, 1, ""), trailing unit
// axes dropped down to ndim0 axes (IO.hpp:121-138, 380-397)
void fill_hypercube(const std::map<std::string, std::string>& kv, int ndim0, fwi2d_hypercube* h)
{
    std::memset(h, 0, sizeof(*h));
    for (int a = 0; a < 6; a++) {
        const std::string k = std::to_string(a + 1);
        auto get = [&](const std::string& key) -> const std::string* {
            auto it = kv.find(key);
            return it == kv.end() ? nullptr : &it->second;
        };
        const std::string* v;
        h->n[a] = (v = get("n" + k)) ? to_ll("n" + k, *v) : 1;
        h->o[a] = (v = get("o" + k)) ? to_d("o" + k, *v) : 0.0;
        h->d[a] = (v = get("d" + k)) ? to_d("d" + k, *v) : 1.0;
        if ((v = get("label" + k))) copy_str(h->label[a], sizeof(h->label[a]), *v);
        if (h->n[a] < 1) throw std::runtime_error("header axis n" + k + " must be at least 1");
    }
    int nd = 6;
    while (nd > 1 && h->n[nd - 1] == 1 && nd != ndim0) nd--;
    h->ndim = nd;
}

long long n123_of(const fwi2d_hypercube* h)
{
    long long n = 1;
    for (int a = 0; a < 6; a++) n *= h->n[a];
    return n;
}

// SEPlib's xdr conversion works on 4-byte words (srite/sreed xdr_float, sep/srite.c:
// 309-360): on this little-endian host each word is byte-swapped
void swap_words(unsigned char* p, size_t nbytes)
{
    for (size_t i = 0; i + 3 < nbytes; i += 4) {
        std::swap(p[i], p[i + 3]);
        std::swap(p[i + 1], p[i + 2]);
    }
}

void read_header(const char* path, int format, int ndim0, fwi2d_hypercube* h)
{
    if (format != FWI2D_FORMAT_SEP && format != FWI2D_FORMAT_BINARY) throw std::runtime_error("unknown file format");
    const std::string p(path);
    const auto kv = parse_pairs(p, format == FWI2D_FORMAT_BINARY);
    fill_hypercube(kv, ndim0, h);
    auto it = kv.find("esize");
    h->esize = it == kv.end() ? 4 : (int)to_ll("esize", it->second);
    if (h->esize != 4 && h->esize != 8)
        throw std::runtime_error("Cannot read into the current template vector: esize " + std::to_string(h->esize));
    h->xdr = 0;
    if (format == FWI2D_FORMAT_SEP) {
        it = kv.find("data_format");
        if (it != kv.end()) {
            const std::string f = strip_quotes(it->second);
            if (f.compare(0, 4, "xdr_") == 0) h->xdr = 1;
            else if (f.compare(0, 6, "native") != 0)
                throw std::runtime_error("unsupported SEPlib data_format " + f);
        }
    }
    it = kv.find("in");
    if (it == kv.end()) throw std::runtime_error("header " + p + " names no binary (in=)");
    std::string bin = strip_quotes(it->second);
    // a relative binary path is taken from the working directory, else next to the header
    if (!bin.empty() && bin[0] != '/' && !file_exists(bin)) {
        const std::string alt = dirname_of(p) + "/" + bin;
        if (file_exists(alt)) bin = alt;
    }
    if (bin.size() >= sizeof(h->in)) throw std::runtime_error("binary path too long: " + bin);
    copy_str(h->in, sizeof(h->in), bin);
}

void read_data(const fwi2d_hypercube* h, float* out)
{
    const long long n = n123_of(h);
    const int es = h->esize;
    FILE* f = std::fopen(h->in, "rb");
    if (!f) throw std::runtime_error(std::string("The input file ") + h->in + " could not be read");
    std::fseek(f, 0, SEEK_END);
    const long long size = std::ftell(f);
    std::rewind(f);
    if (size != n * es) {
        std::fclose(f);
        throw std::runtime_error("The binary file contains " + std::to_string(size) +
                                 " bytes whereas the description file suggests " + std::to_string(n * es) + " bytes");
    }
    const long long chunk = 1 << 20;
    std::vector<unsigned char> buf((size_t)std::min(n, chunk) * es);
    for (long long i0 = 0; i0 < n; i0 += chunk) {
        const long long m = std::min(chunk, n - i0);
        if ((long long)std::fread(buf.data(), es, (size_t)m, f) != m) {
            std::fclose(f);
            throw std::runtime_error(std::string("The binary file ") + h->in + " was not read properly");
        }
        if (h->xdr) swap_words(buf.data(), (size_t)(m * es));
        if (es == 4) std::memcpy(out + i0, buf.data(), (size_t)m * 4);
        else
            for (long long i = 0; i < m; i++) {
                double v;
                std::memcpy(&v, buf.data() + i * 8, 8);
                out[i0 + i] = (float)v;
            }
    }
    std::fclose(f);
}

std::string format_double(double v)
{
    char b[64];
    std::snprintf(b, sizeof(b), "%.9g", v);
    return b;
}

void write_file(const char* path, int format, const fwi2d_hypercube* h, const float* data, const char* datapath)
{
    if (format != FWI2D_FORMAT_SEP && format != FWI2D_FORMAT_BINARY) throw std::runtime_error("unknown file format");
    if (h->ndim < 1 || h->ndim > 6) throw std::runtime_error("the hypercube needs 1 to 6 axes");
    for (int a = 0; a < h->ndim; a++)
        if (h->n[a] < 1) throw std::runtime_error("axis lengths must be at least 1");
    const std::string p(path);
    // the binary: datapath/<header name>@, datapath from the argument, else $DATAPATH,
    // else the header's own directory
    std::string dp = datapath && *datapath ? std::string(datapath) : std::string();
    if (dp.empty()) {
        const char* env = std::getenv("DATAPATH");
        dp = env && *env ? std::string(env) : dirname_of(p);
    }
    if (dp.back() != '/') dp += "/";
    const std::string bin = dp + basename_of(p) + "@";
    long long n = 1;
    for (int a = 0; a < h->ndim; a++) n *= h->n[a];
    {
        std::ofstream hf(p);
        if (!hf.is_open()) throw std::runtime_error("The output file " + p + " could not be created");
        for (int a = 0; a < h->ndim; a++) {
            const std::string k = std::to_string(a + 1);
            if (format == FWI2D_FORMAT_SEP) {
                hf << "\t\tn" << k << "=" << h->n[a] << " o" << k << "=" << format_double(h->o[a]) << " d" << k
                   << "=" << format_double(h->d[a]);
                if (h->label[a][0]) hf << " label" << k << "=\"" << h->label[a] << "\"";
                hf << "\n";
            } else {
                hf << "n" << k << "=" << h->n[a] << "\n"
                   << "o" << k << "=" << format_double(h->o[a]) << "\n"
                   << "d" << k << "=" << format_double(h->d[a]) << "\n";
                if (h->label[a][0]) hf << "label" << k << "=" << h->label[a] << "\n";
            }
        }
        if (format == FWI2D_FORMAT_SEP) hf << "\t\tesize=4 data_format=\"xdr_float\"\n\t\tin=\"" << bin << "\"\n";
        else hf << "esize=4\nin=" << bin << "\n";
        if (!hf.good()) throw std::runtime_error("could not write the header " + p);
    }
    FILE* f = std::fopen(bin.c_str(), "wb");
    if (!f) throw std::runtime_error("The output file " + bin + " could not be created. Check the path");
    const long long chunk = 1 << 20;
    std::vector<unsigned char> buf((size_t)std::min(n, chunk) * 4);
    for (long long i0 = 0; i0 < n; i0 += chunk) {
        const long long m = std::min(chunk, n - i0);
        std::memcpy(buf.data(), data + i0, (size_t)m * 4);
        if (format == FWI2D_FORMAT_SEP) swap_words(buf.data(), (size_t)m * 4);
        if ((long long)std::fwrite(buf.data(), 4, (size_t)m, f) != m) {
            std::fclose(f);
            throw std::runtime_error("Cannot write data to " + bin);
        }
    }
    if (std::fclose(f) != 0) throw std::runtime_error("Cannot write data to " + bin);
}

struct Coord {
    std::vector<double> sxz;
    std::vector<int> nrec;
    std::vector<double> rxz;  // (x, z, dip) per receiver
};

// readCoord (param.cpp:4-77): a line that does not start with '#' holds a shot id, the
// shot (x, z) and one receiver (x, z, dip); consecutive lines with the same id (its
// integer part) add receivers to the shot, a new id starts a new shot.  Values are read
// as the reference's data_t (single precision).  Missing trailing values read as 0.
Coord read_coord(const char* path)
{
    std::ifstream in(path);
    if (!in.is_open()) throw std::runtime_error(std::string("The geometry file ") + path + " could not be opened");
    Coord c;
    std::string line;
    int prev = 0;
    bool first = true;
    while (std::getline(in, line)) {
        if (!line.empty() && line[0] == '#') continue;
        std::istringstream ss(line);
        float sid;
        if (!(ss >> sid)) continue;  // a blank line
        const int cur = (int)sid;
        float v[5] = {0, 0, 0, 0, 0};
        for (float& x : v)
            if (!(ss >> x)) break;
        if (first || cur != prev) {
            c.sxz.push_back(v[0]);
            c.sxz.push_back(v[1]);
            c.nrec.push_back(0);
        }
        c.rxz.push_back(v[2]);
        c.rxz.push_back(v[3]);
        c.rxz.push_back(v[4]);
        c.nrec.back()++;
        prev = cur;
        first = false;
    }
    return c;
}

}  // namespace

extern "C" {

int fwi2d_read_header(const char* path, int format, int ndim0, fwi2d_hypercube* h)
{
    return io_guard([&] {
        if (!path || !h) throw std::runtime_error("null argument");
        read_header(path, format, ndim0, h);
    });
}

int fwi2d_read_data(const fwi2d_hypercube* h, float* data)
{
    return io_guard([&] {
        if (!h || !data) throw std::runtime_error("null argument");
        read_data(h, data);
    });
}

int fwi2d_write(const char* path, int format, const fwi2d_hypercube* h, const float* data, const char* datapath)
{
    return io_guard([&] {
        if (!path || !h || !data) throw std::runtime_error("null argument");
        write_file(path, format, h, data, datapath);
    });
}

int fwi2d_read_coord(const char* path, int* ns, long long* nrec_total, double* sxz, int* nrec, double* rxz)
{
    return io_guard([&] {
        if (!path || !ns || !nrec_total) throw std::runtime_error("null argument");
        const Coord c = read_coord(path);
        const int n = (int)c.nrec.size();
        const long long nr = (long long)c.rxz.size() / 3;
        if (sxz || nrec || rxz) {
            if (*ns < n || *nrec_total < nr) throw std::runtime_error("geometry arrays too small");
            if (sxz) std::copy(c.sxz.begin(), c.sxz.end(), sxz);
            if (nrec) std::copy(c.nrec.begin(), c.nrec.end(), nrec);
            if (rxz) std::copy(c.rxz.begin(), c.rxz.end(), rxz);
        }
        *ns = n;
        *nrec_total = nr;
    });
}

int fwi2d_regular_geometry(int ns, float sx0, float sz0, float sxinc, float szinc, int nr, float rx0, float rz0,
                           float rxinc, float rzinc, float rdip0, float rdipinc, double* sxz, double* rxz)
{
    return io_guard([&] {
        if (ns < 0 || nr < 0) throw std::runtime_error("negative shot or receiver count");
        if ((ns && !sxz) || (ns && nr && !rxz)) throw std::runtime_error("null argument");
        for (int s = 0; s < ns; s++) {
            // data_t arithmetic, as param.cpp:233 and 238
            sxz[2 * s] = (float)(sx0 + (float)s * sxinc);
            sxz[2 * s + 1] = (float)(sz0 + (float)s * szinc);
            for (int r = 0; r < nr; r++) {
                double* q = rxz + 3 * ((long long)s * nr + r);
                q[0] = (float)(rx0 + (float)r * rxinc);
                q[1] = (float)(rz0 + (float)r * rzinc);
                q[2] = (float)(rdip0 + (float)r * rdipinc);
            }
        }
    });
}

}  // extern "C"
