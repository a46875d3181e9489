/*
 * Hdf5.hpp -- HDF5 reading and writing without libhdf5, for the files the
 * reference reads and writes (hdfHandler.cpp:12-254): a group of one- and
 * two-dimensional numeric datasets.
 *
 * The reader follows the HDF5 file format specification for what such files
 * hold: superblock version 0 or 1 with 8-byte offsets and lengths, version 1
 * object headers (with continuation blocks), symbol-table groups (version 1
 * B-trees, symbol table nodes, local heaps), contiguous or compact storage
 * without filters, little-endian IEEE floats and two's-complement integers.
 * Anything else is reported as unsupported, never guessed at.  Values are
 * converted as H5Dread converts them to H5T_NATIVE_FLOAT / H5T_NATIVE_INT32:
 * doubles round to nearest, integers outside int32 clamp.
 *
 * The writer produces the same subset (superblock 0, one symbol-table node
 * per group, contiguous datasets), which libhdf5 and h5dump read.
 */
#ifndef SPH_HOST_HDF5_HPP
#define SPH_HOST_HDF5_HPP

#include <cstdint>
#include <map>
#include <string>
#include <vector>

namespace h5 {

class Reader {
public:
    /* Reads the file into memory and parses the superblock. */
    bool open(const std::string& path);
    const std::string& error() const { return err_; }
    /* "Group/Dataset" (any depth).  false when absent or unsupported. */
    bool readFloat(const std::string& path, std::vector<float>& out, std::vector<uint64_t>& dims);
    bool readInt32(const std::string& path, std::vector<int32_t>& out, std::vector<uint64_t>& dims);
    bool exists(const std::string& path);

private:
    struct Header {
        uint64_t btree = 0, heap = 0; /* symbol table message */
        bool group = false;
        std::vector<uint64_t> dims;
        bool have_space = false, have_type = false, have_layout = false, filtered = false, external = false;
        int type_class = -1, type_size = 0, byte_order = 0;
        bool is_signed = false;
        int layout_class = -1; /* 0 compact, 1 contiguous, 2 chunked */
        uint64_t data_addr = 0, data_size = 0;
        std::vector<uint8_t> compact;
    };
    bool fail(const std::string& msg);
    bool range(uint64_t addr, uint64_t len) const;
    uint64_t u(uint64_t addr, int bytes) const;
    bool parseHeader(uint64_t addr, Header& h);
    bool parseMessage(int type, const uint8_t* p, uint64_t len, Header& h, std::vector<std::pair<uint64_t, uint64_t>>& more);
    bool groupEntries(const Header& g, std::map<std::string, uint64_t>& out);
    bool btreeEntries(uint64_t node, uint64_t heap_data, uint64_t heap_size, std::map<std::string, uint64_t>& out,
                      int depth);
    bool locate(const std::string& path, Header& h);
    bool raw(const std::string& path, Header& h, const uint8_t*& data, uint64_t& count);

    std::vector<uint8_t> buf_;
    uint64_t base_ = 0, root_ = 0;
    uint64_t budget_ = 0; /* B-tree nodes and symbols left to visit (a corrupt file may loop) */
    std::string err_;
};

class Writer {
public:
    /* Datasets are added to groups below the root ("PartType0"); dims of rank
     * 1 or 2.  The data is copied. */
    void addFloat(const std::string& group, const std::string& name, const std::vector<uint64_t>& dims,
                  const float* data);
    void addInt32(const std::string& group, const std::string& name, const std::vector<uint64_t>& dims,
                  const int32_t* data);
    bool write(const std::string& path, std::string& err) const;

private:
    struct Dataset {
        std::string name;
        std::vector<uint64_t> dims;
        bool is_float = true;
        std::vector<uint8_t> bytes;
    };
    std::map<std::string, std::vector<Dataset>> groups_;
};

}  // namespace h5

#endif
