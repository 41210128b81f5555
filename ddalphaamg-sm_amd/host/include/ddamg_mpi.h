// MPI for the host API.  One process drives one GPU that holds the whole lattice, so the
// library itself makes no MPI call.  A caller written against the reference (its main
// calls MPI_Init, MPI_Bcast, MPI_Wtime, ...) gets the real <mpi.h> when the compiler
// finds one, and otherwise the single-process definitions below: rank 0 of a world of
// size 1, broadcasts and reductions that leave or copy the data, wall-clock MPI_Wtime.
#ifndef DDAMG_MPI_H
#define DDAMG_MPI_H

#if defined(__has_include)
#if __has_include(<mpi.h>)
#define DDAMG_HAVE_MPI 1
#endif
#endif

#ifdef DDAMG_HAVE_MPI
#include <mpi.h>
#else
#include <chrono>
#include <cstring>

typedef int MPI_Comm;
typedef int MPI_Datatype;
typedef int MPI_Op;
#define MPI_COMM_WORLD 0
#define MPI_COMM_NULL (-1)
#define MPI_CHAR 1
#define MPI_INT 4
#define MPI_LONG_LONG_INT 8
#define MPI_DOUBLE 8
#define MPI_DOUBLE_COMPLEX 16
#define MPI_SUM 0
#define MPI_SUCCESS 0

// internal linkage: no copy of these leaks into a program linked with a real MPI
static inline int MPI_Init(int*, char***) { return MPI_SUCCESS; }
static inline int MPI_Finalize() { return MPI_SUCCESS; }
static inline int MPI_Comm_size(MPI_Comm, int* size) {
    *size = 1;
    return MPI_SUCCESS;
}
static inline int MPI_Comm_rank(MPI_Comm, int* rank) {
    *rank = 0;
    return MPI_SUCCESS;
}
static inline int MPI_Barrier(MPI_Comm) { return MPI_SUCCESS; }
static inline int MPI_Bcast(void*, int, MPI_Datatype, int, MPI_Comm) { return MPI_SUCCESS; }
// the datatype constants above are the element sizes in bytes
static inline int MPI_Allreduce(const void* in, void* out, int count, MPI_Datatype type, MPI_Op, MPI_Comm) {
    if (in != out) std::memmove(out, in, (size_t)count * (size_t)type);
    return MPI_SUCCESS;
}
static inline double MPI_Wtime() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}
#endif

#endif
