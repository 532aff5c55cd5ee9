/*
 * minimc_amd_h5.h — pandas-HDF5 tables (host library libmmc_h5.so).
 *
 * The reference reads every continuous-energy and thermal-scattering table
 * through HDF5DataSet<D> (src/HDF5DataSet.hpp:18-167): a pandas DataFrame
 * with one column and a D-level MultiIndex, stored under the key "pandas"
 * (group /pandas, attribute axis1_nlevels = D, datasets axis1_level<i> and
 * block0_values, last level fastest).  These entry points read and write
 * that layout; the World builder turns the tables into mmc_table /
 * mmc_tsl descriptors.  Messages (mmc_h5_last_error) follow the reference's
 * where it has one ("File not found: <path>", ReadPandasAxis :89-93).
 */
#ifndef MINIMC_AMD_H5_H
#define MINIMC_AMD_H5_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* HDF5DataSet::ReadPandasAxis / ReadPandasValues (HDF5DataSet.hpp:87-130):
 * the level count, the size of each level (first max_levels) and the number
 * of values.  Returns 0, or 1 with a message. */
int mmc_h5_info(const char* path, int32_t* levels, uint64_t* sizes,
                int32_t max_levels, uint64_t* n_values);
/* The levels (axes[i] holds sizes[i] doubles) and the values. */
int mmc_h5_read(const char* path, int32_t levels, double* const* axes,
                double* values);
/* Writes a new file in the same layout (no counterpart in the reference,
 * which reads files pandas wrote): the product of the levels' sizes values,
 * last level fastest. */
int mmc_h5_write(const char* path, int32_t levels, const uint64_t* sizes,
                 const double* const* axes, const double* values);
const char* mmc_h5_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* MINIMC_AMD_H5_H */
