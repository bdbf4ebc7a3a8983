// Read access to the NetCDF files CENTRE consumes, whatever their format:
// the classic family (CDF-1/2/5, read by nc_classic) and NetCDF-4, which is
// HDF5 underneath (variables are root-group datasets of the same name, global
// attributes are root-group attributes).  NetCDF-4 reads load libhdf5 at run
// time (dlopen); without it such files are refused with a clear error.
//
// Replaces the read side of NetcdfFile / Netcdf_BAT / Netcdf_TrjInt /
// Netcdf_EntContri (NetcdfFile.cpp:16-110, Netcdf_BAT.cpp:24-132,
// Netcdf_TrjInt.cpp:176-370, Netcdf_EntContri.cpp:173-420), which go through
// netcdf-c.
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "nc_classic.hpp"

namespace ncf {

class File {
public:
    virtual ~File() = default;
    virtual bool has_var(const std::string& name) const = 0;
    virtual std::vector<uint64_t> shape(const std::string& name) const = 0;
    // Hyperslab reads (empty start/count = whole variable), converted to the
    // requested element type.
    virtual std::vector<double> read_f64(const std::string& name, std::vector<uint64_t> start = {},
                                         std::vector<uint64_t> count = {}) const = 0;
    virtual std::vector<uint8_t> read_u8(const std::string& name, std::vector<uint64_t> start = {},
                                         std::vector<uint64_t> count = {}) const = 0;
    virtual std::vector<uint32_t> read_u32(const std::string& name, std::vector<uint64_t> start = {},
                                           std::vector<uint64_t> count = {}) const = 0;
    // Text of a global attribute, "" when absent or not text.
    virtual std::string text_attr(const std::string& name) const = 0;
    // Values of an integer global attribute, empty when absent or not integer.
    virtual std::vector<uint64_t> uint_attr(const std::string& name) const = 0;
    virtual const char* format() const = 0;
};

// Opens by content (magic bytes), not by file name.
std::unique_ptr<File> open(const std::string& path);

// Writes the file a classic-format Writer describes as NetCDF-4 (HDF5 with
// netcdf-c's conventions: one dimension scale per dimension carrying
// _Netcdf4Dimid, DIMENSION_LIST / REFERENCE_LIST links, _Netcdf4Coordinates
// on variables of rank > 1, creation order tracked, NetCDF default fill
// values, _NCProperties), the way the reference creates its
// entropy-contribution files (Netcdf_EntContri.cpp:577-760, NC_NETCDF4).
// Variables with the unlimited dimension are chunked (1, full extent, ...)
// unless the Writer names their chunks.  Loads libhdf5 and libhdf5_hl at run
// time; throws nc::Error when they are absent.
void write_netcdf4(const nc::Writer& w, const std::string& path);
// True when the libraries write_netcdf4 needs can be loaded.
bool netcdf4_available();

}  // namespace ncf
