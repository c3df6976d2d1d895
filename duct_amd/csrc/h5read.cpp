// The HDF5 reader of the Fourier-space field files. Format: HDF5 File Format Specification
// (superblock, object header, local heap, v1 B-tree and symbol-table node layouts, and the
// dataspace 0x0001, datatype 0x0003, data layout 0x0008, link 0x0006, link info 0x0002,
// continuation 0x0010 and symbol table 0x0011 messages).
#include "h5read.h"

#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace duct_h5 {

namespace {

[[noreturn]] void bad(const std::string& path, const std::string& what)
{
    throw std::runtime_error(path + ": " + what);
}

bool sig(const uint8_t* p, const char* s) { return std::memcmp(p, s, 4) == 0; }

uint64_t le(const uint8_t* p, int n)
{
    uint64_t v = 0;
    for (int i = n - 1; i >= 0; --i) v = (v << 8) | p[i];
    return v;
}

constexpr int kMaxChunks = 4096;    // continuation chunks followed per object header
constexpr int kMaxDepth = 32;       // B-tree levels

}  // namespace

uint64_t Dataset::count() const
{
    uint64_t c = 1;
    for (uint64_t d : dims) c *= d;
    return c;
}

File::File(const std::string& path) : path_(path)
{
    fd_ = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd_ < 0) bad(path, std::string("cannot open: ") + std::strerror(errno));
    struct stat st{};
    if (::fstat(fd_, &st) != 0) {
        ::close(fd_);
        bad(path, "cannot stat");
    }
    file_size_ = (uint64_t)st.st_size;
    try {
        read_root();
    } catch (...) {
        ::close(fd_);
        fd_ = -1;
        throw;
    }
}

File::~File()
{
    if (fd_ >= 0) ::close(fd_);
}

void File::pread_exact(uint64_t off, void* dst, size_t n) const
{
    if (off > file_size_ || n > file_size_ - off) bad(path_, "read past the end of the file (truncated or corrupt)");
    uint8_t* p = (uint8_t*)dst;
    while (n) {
        const ssize_t r = ::pread(fd_, p, n, (off_t)off);
        if (r < 0 && errno == EINTR) continue;
        if (r <= 0) bad(path_, "read error");
        p += r; off += (uint64_t)r; n -= (size_t)r;
    }
}

uint64_t File::rd_uint(uint64_t off, int nbytes) const
{
    uint8_t b[8];
    pread_exact(off, b, (size_t)nbytes);
    return le(b, nbytes);
}

bool File::undefined(uint64_t a) const
{
    return so_ == 8 ? a == ~0ull : a == ((1ull << (8 * so_)) - 1);
}

// The superblock is at 0, 512, 1024, 2048, ... (after a user block).
void File::read_root()
{
    uint64_t sb = 0;
    for (;; sb = sb ? 2 * sb : 512) {
        if (sb + 8 > file_size_) bad(path_, "no HDF5 superblock (not an HDF5 file)");
        uint8_t s[8];
        pread_exact(sb, s, 8);
        if (std::memcmp(s, "\x89HDF\r\n\x1a\n", 8) == 0) break;
    }
    uint8_t h[16];
    pread_exact(sb, h, 16);
    const int ver = h[8];
    uint64_t root_ohdr = 0;
    if (ver == 0 || ver == 1) {
        so_ = h[13];
        sl_ = h[14];
        if ((so_ != 2 && so_ != 4 && so_ != 8) || (sl_ != 2 && sl_ != 4 && sl_ != 8))
            bad(path_, "unsupported offset or length size in the superblock");
        uint64_t p = sb + 24 + (ver == 1 ? 4 : 0);
        base_ = addr(p);
        p += 4 * (uint64_t)so_;                 // base, free-space, end of file, driver block
        root_ohdr = addr(p + so_);              // root symbol table entry: name offset, header
    } else if (ver == 2 || ver == 3) {
        so_ = h[9];
        sl_ = h[10];
        if ((so_ != 2 && so_ != 4 && so_ != 8) || (sl_ != 2 && sl_ != 4 && sl_ != 8))
            bad(path_, "unsupported offset or length size in the superblock");
        const uint64_t p = sb + 12;
        base_ = addr(p);
        root_ohdr = addr(p + 3 * (uint64_t)so_);  // base, extension, end of file, root header
    } else {
        bad(path_, "unsupported superblock version " + std::to_string(ver));
    }
    bool found = false;
    for (const Msg& m : messages(base_ + root_ohdr)) {
        if (m.type == 0x11) {                    // symbol table: v1 B-tree + local heap
            const uint64_t tree = addr(m.at), heap = addr(m.at + so_);
            uint8_t hh[8];
            pread_exact(base_ + heap, hh, 8);
            if (!sig(hh, "HEAP")) bad(path_, "bad local heap signature");
            const uint64_t heap_data = base_ + addr(base_ + heap + 8 + 2 * (uint64_t)sl_);
            walk_btree(base_ + tree, heap_data, 0);
            found = true;
        } else if (m.type == 0x02) {             // link info: dense storage when the heap is set
            uint8_t v[2];
            pread_exact(m.at, v, 2);
            const uint64_t heap = addr(m.at + 2 + ((v[1] & 1) ? 8 : 0));
            if (!undefined(heap)) bad(path_, "the root group uses dense link storage (unsupported)");
            found = true;
        } else if (m.type == 0x06) {             // link (compact storage)
            std::vector<uint8_t> b(m.size);
            pread_exact(m.at, b.data(), m.size);
            if (m.size < 3 || b[0] != 1) bad(path_, "unsupported link message");
            const int fl = b[1];
            size_t q = 2;
            int ltype = 0;
            if (fl & 0x08) ltype = b[q++];
            if (fl & 0x04) q += 8;
            if (fl & 0x10) q += 1;
            const int nl = 1 << (fl & 3);
            if (q + nl > b.size()) bad(path_, "truncated link message");
            const uint64_t n = le(&b[q], nl);
            q += nl;
            if (q + n > b.size()) bad(path_, "truncated link message");
            std::string name((const char*)&b[q], (size_t)n);
            q += n;
            if (ltype == 0) {
                if (q + so_ > b.size()) bad(path_, "truncated link message");
                links_.emplace_back(name, base_ + le(&b[q], so_));
            }                                     // soft and external links are not followed
            found = true;
        }
    }
    if (!found) bad(path_, "the root object has no group messages");
}

std::string File::heap_name(uint64_t heap_data, uint64_t off) const
{
    std::string s;
    char buf[64];
    for (uint64_t p = heap_data + off;; p += sizeof buf) {
        const size_t n = (size_t)std::min<uint64_t>(sizeof buf, file_size_ - std::min(p, file_size_));
        if (n == 0) bad(path_, "unterminated name in the local heap");
        pread_exact(p, buf, n);
        const void* z = std::memchr(buf, 0, n);
        if (z) return s.append(buf, (const char*)z - buf);
        s.append(buf, n);
        if (s.size() > 4096) bad(path_, "name in the local heap too long");
    }
}

void File::walk_btree(uint64_t node, uint64_t heap_data, int depth)
{
    if (depth > kMaxDepth) bad(path_, "group B-tree too deep (corrupt)");
    uint8_t h[8];
    pread_exact(node, h, 8);
    if (!sig(h, "TREE")) bad(path_, "bad B-tree signature");
    if (h[4] != 0) bad(path_, "the root B-tree is not a group node");
    const int level = h[5];
    const int used = (int)le(h + 6, 2);
    const uint64_t first = node + 8 + 2 * (uint64_t)so_;   // after the two sibling addresses
    for (int i = 0; i < used; ++i) {
        const uint64_t child = base_ + addr(first + (uint64_t)sl_ + (uint64_t)i * (sl_ + so_));
        if (level > 0) {
            walk_btree(child, heap_data, depth + 1);
            continue;
        }
        uint8_t s[8];
        pread_exact(child, s, 8);
        if (!sig(s, "SNOD")) bad(path_, "bad symbol table node signature");
        const int nsym = (int)le(s + 6, 2);
        const uint64_t esz = 2 * (uint64_t)so_ + 24;
        for (int e = 0; e < nsym; ++e) {
            const uint64_t ent = child + 8 + (uint64_t)e * esz;
            links_.emplace_back(heap_name(heap_data, addr(ent)), base_ + addr(ent + so_));
        }
    }
}

std::vector<File::Msg> File::messages(uint64_t ohdr) const
{
    std::vector<Msg> out;
    uint8_t h[16];
    pread_exact(ohdr, h, 16);
    struct Chunk { uint64_t b, e; };
    std::vector<Chunk> todo;
    bool v2 = false, order = false;
    if (sig(h, "OHDR")) {
        v2 = true;
        if (h[4] != 2) bad(path_, "unsupported object header version");
        const int fl = h[5];
        order = fl & 0x04;
        uint64_t p = ohdr + 6 + ((fl & 0x20) ? 16 : 0) + ((fl & 0x10) ? 4 : 0);
        const int nb = 1 << (fl & 3);
        const uint64_t sz = rd_uint(p, nb);
        p += nb;
        todo.push_back({p, p + sz});
    } else {
        if (h[0] != 1) bad(path_, "unsupported object header version");
        const uint64_t sz = le(h + 8, 4);
        todo.push_back({ohdr + 16, ohdr + 16 + sz});
    }
    const int mh = v2 ? (order ? 6 : 4) : 8;     // bytes of a message's own header
    for (size_t c = 0; c < todo.size(); ++c) {
        if ((int)c > kMaxChunks) bad(path_, "too many object header continuations (corrupt)");
        const Chunk ch = todo[c];
        if (ch.e > file_size_ || ch.b > ch.e) bad(path_, "object header outside the file");
        std::vector<uint8_t> b(ch.e - ch.b);
        pread_exact(ch.b, b.data(), b.size());
        size_t q = 0;
        while (q + mh <= b.size()) {
            int type;
            uint32_t size;
            if (v2) {
                type = b[q];
                size = (uint32_t)le(&b[q + 1], 2);
            } else {
                type = (int)le(&b[q], 2);
                size = (uint32_t)le(&b[q + 2], 2);
            }
            const uint64_t at = ch.b + q + mh;
            if (q + mh + size > b.size()) bad(path_, "object header message past its chunk");
            if (type == 0x10) {                  // continuation: offset, length
                const uint64_t off = base_ + addr(at), n = len(at + so_);
                if (v2) {
                    uint8_t s4[4];
                    pread_exact(off, s4, 4);
                    if (!sig(s4, "OCHK") || n < 8) bad(path_, "bad continuation block");
                    todo.push_back({off + 4, off + n - 4});   // signature .. checksum
                } else {
                    todo.push_back({off, off + n});
                }
            } else if (type != 0) {
                out.push_back({type, at, size});
            }
            q += mh + size;
        }
    }
    return out;
}

bool File::has(const std::string& name) const
{
    for (const auto& l : links_)
        if (l.first == name) return true;
    return false;
}

Dataset File::dataset(const std::string& name) const
{
    uint64_t oh = 0;
    bool found = false;
    for (const auto& l : links_)
        if (l.first == name) { oh = l.second; found = true; break; }
    if (!found) bad(path_, "no dataset '" + name + "' in the root group");
    Dataset d;
    bool have_space = false, have_type = false, have_layout = false;
    for (const Msg& m : messages(oh)) {
        std::vector<uint8_t> b(m.size);
        pread_exact(m.at, b.data(), m.size);
        if (m.type == 0x01) {                    // dataspace
            if (m.size < 4) bad(path_, "truncated dataspace message");
            const int ver = b[0], nd = b[1];
            size_t q;
            if (ver == 1) q = 8;
            else if (ver == 2) q = 4;
            else bad(path_, "unsupported dataspace version");
            if (q + (size_t)nd * sl_ > b.size()) bad(path_, "truncated dataspace message");
            for (int i = 0; i < nd; ++i) d.dims.push_back(le(&b[q + (size_t)i * sl_], sl_));
            if (ver == 2 && b[3] == 2) d.dims.assign(1, 0);   // null dataspace: no elements
            have_space = true;
        } else if (m.type == 0x03) {             // datatype
            if (m.size < 8) bad(path_, "truncated datatype message");
            const int cls = b[0] & 0x0f;
            d.size = (int)le(&b[4], 4);
            d.big_endian = b[1] & 1;
            if (cls == 0) { d.kind = Kind::Int; d.is_signed = b[1] & 0x08; }
            else if (cls == 1) {
                d.kind = Kind::Float;
                if (b[1] & 0x40) bad(path_, "VAX-order floating point is not supported");
            } else if (cls == 3) d.kind = Kind::String;
            else d.kind = Kind::Other;
            have_type = true;
        } else if (m.type == 0x08) {             // data layout
            if (m.size < 2) bad(path_, "truncated layout message");
            const int ver = b[0];
            if (ver == 1 || ver == 2) {
                if (b.size() < 8) bad(path_, "truncated layout message");
                const int nd = b[1], cls = b[2];
                size_t q = 8;
                if (cls == 1) {
                    if (q + (size_t)so_ > b.size()) bad(path_, "truncated layout message");
                    d.addr = le(&b[q], so_);
                    d.compact = false;
                } else if (cls == 0) {
                    q += (size_t)nd * 4;
                    if (q + 4 > b.size()) bad(path_, "truncated layout message");
                    const uint64_t n = le(&b[q], 4);
                    if (q + 4 + n > b.size()) bad(path_, "truncated layout message");
                    d.inline_data.assign(b.begin() + (long)q + 4, b.begin() + (long)(q + 4 + n));
                    d.compact = true;
                } else {
                    bad(path_, "dataset '" + name + "' is chunked (unsupported)");
                }
            } else if (ver == 3 || ver == 4) {
                const int cls = b[1];
                if (cls == 0) {
                    if (b.size() < 4) bad(path_, "truncated layout message");
                    const uint64_t n = le(&b[2], 2);
                    if (4 + n > b.size()) bad(path_, "truncated layout message");
                    d.inline_data.assign(b.begin() + 4, b.begin() + 4 + (long)n);
                    d.compact = true;
                } else if (cls == 1) {
                    if (2 + (size_t)so_ + sl_ > b.size()) bad(path_, "truncated layout message");
                    d.addr = le(&b[2], so_);
                    d.bytes = le(&b[2 + so_], sl_);
                    d.compact = false;
                } else {
                    bad(path_, "dataset '" + name + "' is chunked or virtual (unsupported)");
                }
            } else {
                bad(path_, "unsupported layout version");
            }
            have_layout = true;
        } else if (m.type == 0x11 || m.type == 0x02 || m.type == 0x06) {
            bad(path_, "'" + name + "' is a group, not a dataset");
        }
    }
    if (!have_space || !have_type || !have_layout) bad(path_, "'" + name + "' is not a dataset");
    uint64_t total = (uint64_t)d.size;
    for (uint64_t e : d.dims)
        if (__builtin_mul_overflow(total, e, &total)) bad(path_, "dataset '" + name + "' extents overflow");
    if (!d.compact) {
        if (undefined(d.addr)) bad(path_, "dataset '" + name + "' has no storage (never written)");
        d.addr += base_;
        const uint64_t need = d.count() * (uint64_t)d.size;
        if (d.bytes == 0) d.bytes = need;
        if (d.bytes < need || d.addr > file_size_ || need > file_size_ - d.addr)
            bad(path_, "dataset '" + name + "' extends past the end of the file");
    } else if (d.inline_data.size() < d.count() * (uint64_t)d.size) {
        bad(path_, "compact dataset '" + name + "' is shorter than its dataspace");
    }
    return d;
}

void File::read_doubles(const Dataset& d, uint64_t first, uint64_t n, double* out) const
{
    if (d.kind != Kind::Float && d.kind != Kind::Int) bad(path_, "dataset is not numeric");
    if (!(d.size == 1 || d.size == 2 || d.size == 4 || d.size == 8) || (d.kind == Kind::Float && d.size < 4))
        bad(path_, "unsupported element size " + std::to_string(d.size));
    if (first > d.count() || n > d.count() - first) bad(path_, "element range outside the dataset");
    const int es = d.size;
    if (d.kind == Kind::Float && es == 8) {      // the fields: straight into out, swapped in place
        const uint64_t off = first * 8;
        if (d.compact) std::memcpy(out, d.inline_data.data() + off, (size_t)(n * 8));
        else pread_exact(d.addr + off, out, (size_t)(n * 8));
        if (d.big_endian) {
            uint64_t* u = reinterpret_cast<uint64_t*>(out);
            for (uint64_t i = 0; i < n; ++i) u[i] = __builtin_bswap64(u[i]);
        }
        return;
    }
    constexpr uint64_t kBlock = 1 << 17;         // elements per read
    std::vector<uint8_t> buf((size_t)std::min(n, kBlock) * es);
    for (uint64_t done = 0; done < n;) {
        const uint64_t k = std::min(n - done, kBlock);
        const uint64_t off = (first + done) * es;
        if (d.compact) std::memcpy(buf.data(), d.inline_data.data() + off, (size_t)(k * es));
        else pread_exact(d.addr + off, buf.data(), (size_t)(k * es));
        for (uint64_t i = 0; i < k; ++i) {
            uint8_t e[8];
            const uint8_t* s = buf.data() + i * es;
            for (int j = 0; j < es; ++j) e[j] = d.big_endian ? s[es - 1 - j] : s[j];   // to LE
            double v;
            if (d.kind == Kind::Float) {
                if (es == 8) { uint64_t u = le(e, 8); std::memcpy(&v, &u, 8); }
                else { uint32_t u = (uint32_t)le(e, 4); float f; std::memcpy(&f, &u, 4); v = f; }
            } else {
                const uint64_t u = le(e, es);
                if (d.is_signed && es < 8 && (u >> (8 * es - 1)) & 1) v = (double)(int64_t)(u | (~0ull << (8 * es)));
                else if (d.is_signed) v = (double)(int64_t)u;
                else v = (double)u;
            }
            out[done + i] = v;
        }
        done += k;
    }
}

std::string File::read_string(const Dataset& d) const
{
    if (d.kind != Kind::String) bad(path_, "dataset is not a string");
    const uint64_t n = d.count() * (uint64_t)d.size;
    if (n > (1u << 20)) bad(path_, "string dataset too long");
    std::string s((size_t)n, '\0');
    if (d.compact) std::memcpy(&s[0], d.inline_data.data(), (size_t)n);
    else pread_exact(d.addr, &s[0], (size_t)n);
    return s;
}

std::string fortran_string(const char* s, size_t n)
{
    std::string r(s, n);
    const size_t z = r.find('\0');
    if (z != std::string::npos) r.resize(z);
    while (!r.empty() && r.back() == ' ') r.pop_back();
    return r;
}

void cheb_interp(int myr, int mzr, const double* in, int my, int mz, double* out)
{
    const int ky = std::min(20, myr), kz = std::min(20, mzr);
    const double pi = 4.0 * std::atan(1.0);
    const int ny = myr - 1, nz = mzr - 1;
    // cos(pi j k / N) for j <= N, k < K, from (j k) mod 2N so large j k lose nothing to the argument
    auto dct_table = [pi](int nn, int kmax) {
        std::vector<double> t((size_t)(nn + 1) * kmax);
        for (int j = 0; j <= nn; ++j)
            for (int k = 0; k < kmax; ++k) t[(size_t)j * kmax + k] = std::cos(pi * (double)((long)j * k % (2L * nn)) / nn);
        return t;
    };
    const std::vector<double> cy = dct_table(ny, ky), cz = dct_table(nz, kz);
    // y: c1[z][k][p] = s_k (f_0 + (-1)^k f_N + 2 sum_{0<j<N} f_j cos(pi j k / N))
    std::vector<double> c1((size_t)mzr * ky * 2), c2((size_t)kz * ky * 2);
    for (int z = 0; z < mzr; ++z) {
        const double* f = in + (size_t)z * myr * 2;
        double inner[20][2] = {};
        for (int j = 1; j < ny; ++j) {
            const double f0 = f[(size_t)j * 2], f1 = f[(size_t)j * 2 + 1];
            const double* c = &cy[(size_t)j * ky];
            for (int k = 0; k < ky; ++k) {
                inner[k][0] += f0 * c[k];
                inner[k][1] += f1 * c[k];
            }
        }
        for (int k = 0; k < ky; ++k) {
            const double s = (k == 0 || k == ny) ? 0.5 / ny : 1.0 / ny;
            for (int p = 0; p < 2; ++p) {
                const double acc = f[p] + ((k & 1) ? -f[(size_t)ny * 2 + p] : f[(size_t)ny * 2 + p]);
                c1[((size_t)z * ky + k) * 2 + p] = s * (acc + 2.0 * inner[k][p]);
            }
        }
    }
    // z: c2[l][k][p], the same transform along z of the first ky y-coefficients
    for (int l = 0; l < kz; ++l) {
        const double s = (l == 0 || l == nz) ? 0.5 / nz : 1.0 / nz;
        for (int k = 0; k < ky; ++k)
            for (int p = 0; p < 2; ++p) {
                auto g = [&](int z) { return c1[((size_t)z * ky + k) * 2 + p]; };
                double acc = g(0) + ((l & 1) ? -g(nz) : g(nz));
                double inner = 0.0;
                for (int z = 1; z < nz; ++z) inner += g(z) * cz[(size_t)z * kz + l];
                c2[((size_t)l * ky + k) * 2 + p] = s * (acc + 2.0 * inner);
            }
    }
    // the new nodes as the reference forms them: y2 = -dglc(-1, 1) = cos(pi i / (my - 1))
    auto basis = [pi](int n, int kmax, std::vector<double>& t) {
        t.assign((size_t)n * kmax, 0.0);
        for (int i = 0; i < n; ++i) {
            const double x = -(-std::cos(i * (pi / (double)(n - 1))));
            const double th = std::acos(x);
            for (int k = 0; k < kmax; ++k) t[(size_t)i * kmax + k] = std::cos((double)k * th);
        }
    };
    std::vector<double> ty, tz;
    basis(my, ky, ty);
    basis(mz, kz, tz);
    // out(p, i, j) = sum_l (sum_k c(k, l) T_k(y_i)) T_l(z_j)
    std::vector<double> t1((size_t)my * kz * 2);
    for (int i = 0; i < my; ++i)
        for (int l = 0; l < kz; ++l)
            for (int p = 0; p < 2; ++p) {
                double acc = 0.0;
                for (int k = 0; k < ky; ++k) acc += c2[((size_t)l * ky + k) * 2 + p] * ty[(size_t)i * ky + k];
                t1[((size_t)i * kz + l) * 2 + p] = acc;
            }
    for (int j = 0; j < mz; ++j)
        for (int i = 0; i < my; ++i)
            for (int p = 0; p < 2; ++p) {
                double acc = 0.0;
                for (int l = 0; l < kz; ++l) acc += t1[((size_t)i * kz + l) * 2 + p] * tz[(size_t)j * kz + l];
                out[((size_t)j * my + i) * 2 + p] = acc;
            }
}

}  // namespace duct_h5
