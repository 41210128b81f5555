// Binary records of gauge fields / right-hand sides (src/io.cpp:4-55): per (x, t, mu)
// int x, int t, int mu, double re, double im, x outer.
#ifndef IO_H
#define IO_H
#include <string>

#include "variables.h"

void read_binary(const std::string& name, const spinor& U);
void write_binary(const std::string& name, const spinor& U);

#endif
