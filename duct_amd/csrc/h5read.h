// A reader for the HDF5 files of the Fourier-space fields (uvwp_NNNN.h5, uvwpt_NNNN.h5) and their
// header datasets (readwrite_hdf5.f90:5-197 reads them, :266-381 and :853-918 write them).
//
// It parses the on-disk format directly (no libhdf5): superblock versions 0-3, object headers
// versions 1 and 2, groups held as a symbol table (v1 B-tree + local heap) or as compact link
// messages, and datasets with contiguous or compact storage of fixed-point, floating-point or
// fixed-length string elements in either byte order. Datasets in the root group only: that is
// where the reference puts everything. Chunked or filtered storage, dense link storage and
// external links are reported as unsupported.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace duct_h5 {

enum class Kind { Float, Int, String, Other };

struct Dataset {
    std::vector<uint64_t> dims;        // dataspace extents, slowest first (C order); empty: scalar
    Kind kind = Kind::Other;
    int size = 0;                      // bytes per element
    bool big_endian = false;
    bool is_signed = true;
    bool compact = false;
    uint64_t addr = 0;                 // contiguous storage: absolute file offset
    uint64_t bytes = 0;                // contiguous storage: size in bytes
    std::vector<uint8_t> inline_data;  // compact storage
    uint64_t count() const;            // number of elements
};

class File {
public:
    // Opens path and parses its superblock and root group; throws std::runtime_error.
    explicit File(const std::string& path);
    ~File();
    File(const File&) = delete;
    File& operator=(const File&) = delete;

    bool has(const std::string& name) const;
    // The dataset `name` of the root group; throws if it is missing or not a supported dataset.
    Dataset dataset(const std::string& name) const;
    // Elements [first, first + n) in C order, converted to double (numeric datasets).
    void read_doubles(const Dataset& d, uint64_t first, uint64_t n, double* out) const;
    // The raw bytes of a fixed-length string dataset.
    std::string read_string(const Dataset& d) const;

private:
    struct Msg {
        int type;
        uint64_t at;        // file offset of the message data
        uint32_t size;
    };
    void pread_exact(uint64_t off, void* dst, size_t n) const;
    uint64_t rd_uint(uint64_t off, int nbytes) const;
    uint64_t addr(uint64_t off) const { return rd_uint(off, so_); }
    uint64_t len(uint64_t off) const { return rd_uint(off, sl_); }
    bool undefined(uint64_t a) const;
    std::vector<Msg> messages(uint64_t ohdr) const;
    void read_root();
    void walk_btree(uint64_t node, uint64_t heap_data, int depth);
    std::string heap_name(uint64_t heap_data, uint64_t off) const;

    int fd_ = -1;
    uint64_t file_size_ = 0;
    uint64_t base_ = 0;        // base address: every stored address is relative to it
    int so_ = 8, sl_ = 8;      // sizes of offsets and lengths
    std::vector<std::pair<std::string, uint64_t>> links_;   // root group: name -> object header
    std::string path_;
};

// A Fortran character argument (pointer + hidden length) without its trailing blanks and NULs.
std::string fortran_string(const char* s, size_t n);

// cheb_interp (chebyshev.f90:725-765): the plane in, Fortran (2, myr, mzr), on the Chebyshev-
// Gauss-Lobatto grid of myr x mzr points, onto the grid of my x mz points, out (2, my, mz). The
// forward transform is the reference's DCT-I (tcheb_fftw, tchebw.f:97-111: 1/N inside, 1/(2N) at
// both ends); only its first min(20, myr) x min(20, mzr) coefficients are evaluated at the new
// nodes, T_k(y) = cos(k acos(y)). The reference's test for small coefficients (:755) ends in
// `continue`, so it skips nothing, and nothing is skipped here.
void cheb_interp(int myr, int mzr, const double* in, int my, int mz, double* out);

}  // namespace duct_h5
