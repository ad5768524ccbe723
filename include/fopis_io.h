/*
 * fopis_io.h — HDF5 datasets for the files the reference keeps between runs
 * (libfopis_io.so, built against the HDF5 C library when it is present).
 *
 * The reference writes them with HighFive (a C++ wrapper over the same C
 * library): the capacity-matrix cache resume/objects/<obj>.h5
 * (IO::writeCmatrixData / loadCmatrixData, src/Misc/IO.cpp:28-120), the run
 * state resume/environment.h5 (Environment::saveInfo / loadInfo,
 * src/Misc/Environment.cpp:379-438) and resume/grid_id_<id>.h5
 * (RootGrid::saveResumeData / loadResumeData, RootGridIO.cpp:25-290). This
 * library reads and writes plain datasets; the host facade
 * (fopis::GpuRootGrid) lays them out with the reference's names, types and
 * shapes. It needs no GPU.
 *
 * Every function returns 0 on success and a negative code on failure
 * (fopis_io_last_error() then holds a message). Names are paths inside the
 * file ("Potential/phi"); writing creates the intermediate groups.
 */
#ifndef FOPIS_IO_H
#define FOPIS_IO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fopis_h5 fopis_h5;

/* Element types, as the reference's HighFive calls store them:
 * F64 = createDataSet<double>, U64 = createDataSet<size_t>, I32 = createDataSet<int>;
 * F32 and U8 (float, uint8_t) hold the learned guess's training pairs. */
enum fopis_io_type { FOPIS_IO_F64 = 0, FOPIS_IO_U64 = 1, FOPIS_IO_I32 = 2, FOPIS_IO_F32 = 3, FOPIS_IO_U8 = 4 };

/* mode 0: open an existing file read-only; 1: create or truncate (H5F::Create | H5F::Truncate). */
int fopis_io_open(const char* path, int mode, fopis_h5** out);
int fopis_io_close(fopis_h5* f);
/* A dataset of rank 0 (scalar, HighFive::DataSpace::From(value)) to 3; dims
 * in C order; data is dense C order in the given type. */
int fopis_io_write(fopis_h5* f, const char* name, int type, int rank, const int64_t* dims, const void* data);
/* 1 if the dataset exists, 0 if not. */
int fopis_io_exists(fopis_h5* f, const char* name);
/* Rank and dims (up to 8) of a dataset. */
int fopis_io_info(fopis_h5* f, const char* name, int* rank, int64_t* dims);
/* Reads a whole dataset converted to `type`; count must equal its element count. */
int fopis_io_read(fopis_h5* f, const char* name, int type, int64_t count, void* data);
const char* fopis_io_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
