// Binary records of gauge fields / right-hand sides (src/io.cpp:4-55): per (x, t, mu)
// int x, int t, int mu, double re, double im, x outer.
#ifndef IO_H
#define IO_H
#include <string>

#include "variables.h"

void read_binary(const std::string& name, const spinor& U);
void write_binary(const std::string& name, const spinor& U);
// file name from rank 0 to every process (src/io.cpp:58-70): one process, nothing to send
void broadcast_file_name(std::string& name);
// metadata_<Nx>x<Nt>_Levels<L>_{V,K}cycle.dat with the run's parameters and the
// iteration/FLOP counters (src/io.cpp:72-140)
void writeMetadata(int cycle, double tol);

#endif
