/* Minimal single-rank MPI stand-in for compiling the PCSHELL adapter in tests
 * (no MPI library is installed here).  Test infrastructure only. */
#ifndef MOCK_MPI_H
#define MOCK_MPI_H

typedef int MPI_Comm;
typedef int MPI_Datatype;
typedef int MPI_Op;
#define MPI_COMM_WORLD 0
#define MPI_DOUBLE 1
#define MPI_INT 2
#define MPI_MIN 1
#define MPI_MAX 2
#define MPI_SUCCESS 0

int MPI_Comm_size(MPI_Comm comm, int *size);
int MPI_Reduce(const void *send, void *recv, int count, MPI_Datatype type, MPI_Op op, int root,
               MPI_Comm comm);

#endif
