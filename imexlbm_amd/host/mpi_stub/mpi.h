/*
 * mpi.h -- a one-process MPI for builds of the driver without an MPI
 * library.  It provides the calls the reference's driver, solver object and
 * profiler make (src/main.cpp:10-107, src/lbm.cpp:297-299,
 * src/profile.cpp:110-126) with their one-rank meaning: the world holds one
 * rank, barriers return at once, reductions copy their input, and MPI_Wtime
 * reads a monotonic clock.  A build against a real MPI uses that library's
 * own header instead (the build picks one include path or the other).
 */
#ifndef IMEXLBM_MPI_STUB_H
#define IMEXLBM_MPI_STUB_H

#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#ifdef __cplusplus
extern "C" {
#endif

#define IMEXLBM_MPI_STUB 1

typedef int MPI_Comm;
typedef int MPI_Datatype;
typedef int MPI_Op;

#define MPI_COMM_WORLD ((MPI_Comm)0x44000000)
#define MPI_COMM_SELF ((MPI_Comm)0x44000001)
#define MPI_SUCCESS 0
#define MPI_ERR_ARG 12

#define MPI_INT ((MPI_Datatype)0x4c000405)
#define MPI_LONG ((MPI_Datatype)0x4c000807)
#define MPI_DOUBLE ((MPI_Datatype)0x4c00080b)

#define MPI_SUM ((MPI_Op)0x58000003)
#define MPI_MIN ((MPI_Op)0x58000002)
#define MPI_MAX ((MPI_Op)0x58000001)

#define MPI_IN_PLACE ((void *)-1)

static int imexlbm_mpi_stub_initialized_ = 0;

static inline size_t imexlbm_mpi_stub_size_(MPI_Datatype t)
{
    return t == MPI_INT ? sizeof(int) : (t == MPI_LONG ? sizeof(long) : sizeof(double));
}

static inline int MPI_Init(int *argc, char ***argv)
{
    (void)argc;
    (void)argv;
    imexlbm_mpi_stub_initialized_ = 1;
    return MPI_SUCCESS;
}

static inline int MPI_Initialized(int *flag)
{
    *flag = imexlbm_mpi_stub_initialized_;
    return MPI_SUCCESS;
}

static inline int MPI_Finalize(void)
{
    imexlbm_mpi_stub_initialized_ = 0;
    return MPI_SUCCESS;
}

static inline int MPI_Comm_rank(MPI_Comm comm, int *rank)
{
    (void)comm;
    *rank = 0;
    return MPI_SUCCESS;
}

static inline int MPI_Comm_size(MPI_Comm comm, int *size)
{
    (void)comm;
    *size = 1;
    return MPI_SUCCESS;
}

static inline int MPI_Barrier(MPI_Comm comm)
{
    (void)comm;
    return MPI_SUCCESS;
}

static inline double MPI_Wtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* One rank: the reduction of a single contribution is that contribution. */
static inline int MPI_Allreduce(const void *sendbuf, void *recvbuf, int count, MPI_Datatype type,
                                MPI_Op op, MPI_Comm comm)
{
    (void)op;
    (void)comm;
    if (count < 0) return MPI_ERR_ARG;
    if (sendbuf != MPI_IN_PLACE && sendbuf != recvbuf)
        memcpy(recvbuf, sendbuf, (size_t)count * imexlbm_mpi_stub_size_(type));
    return MPI_SUCCESS;
}

static inline int MPI_Reduce(const void *sendbuf, void *recvbuf, int count, MPI_Datatype type,
                             MPI_Op op, int root, MPI_Comm comm)
{
    (void)root;
    return MPI_Allreduce(sendbuf, recvbuf, count, type, op, comm);
}

static inline int MPI_Abort(MPI_Comm comm, int code)
{
    (void)comm;
    fflush(stdout);
    exit(code);
    return MPI_SUCCESS;
}

#ifdef __cplusplus
}
#endif
#endif
