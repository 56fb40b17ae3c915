/*
 * process.h -- the host application's process object (C++), as declared by
 * the host's include/process.h:1-33. The host defines these members and
 * MPI_Process_get; a plugin only calls them (isMaster, cout, cerr, abort).
 */
#ifndef HDDM_MPI_PROCESS
#define HDDM_MPI_PROCESS

#include <mpi.h>

class GoogleTest;

class MPI_Process
{
	static MPI_Process* GoogleTest(int argc, char* argv[]);

	friend class GoogleTest;

public :
	bool isMaster() const;
	int getRoot() const;
	int getRank() const;
	int getSize() const;
	MPI_Comm getComm() const;
	void setComm(MPI_Comm comm);

	void abort();
	void cout(const char* format, ...);
	void cerr(const char* format, ...);
};

extern "C" int MPI_Process_get(MPI_Process**);

#endif // HDDM_MPI_PROCESS
