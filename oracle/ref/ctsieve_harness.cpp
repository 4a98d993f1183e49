// Test infrastructure: drives the reference CPU window sieve, compiled from its
// own source file (sieve/CPU/ctsieve.cpp), with explicit p bounds instead of the
// hard-coded ini bounds of Sieve::check (ctsieve.cpp:800-801), and writes the
// survivors (b in ascending order, one per line: the cand_*.txt format).
//
// usage: ctsieve_ref m b_min b_max p_max_pos p_max_neg out.txt [dir]
//        ctsieve_ref m b_min b_max p_max_pos p_max_neg out.txt -p p_min_pos p_min_neg
//
// With dir, the reference's own Sieve::write(true) then writes its state file
// sv_<m>_<b_min>_<b_max>.dat and cand_<m>_<b_min>_<b_max>.txt into dir (the
// byte format fixtures of the file tests).  With -p the sieve starts from the
// given p bounds instead of the ini start (the resume state of ctsieve.cpp:790-796).
//
// The standard headers the reference includes are included first so that the
// access override below only applies to the reference's own class.
#include <cstdint>
#include <stdexcept>
#include <iostream>
#include <iomanip>
#include <fstream>
#include <sstream>
#include <string>
#include <cmath>
#include <vector>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <sys/stat.h>
#include <signal.h>
#include <unistd.h>

#define main ctsieve_reference_main
#define private public
#include "ctsieve.cpp"
#undef private
#undef main

int main(int argc, char **argv)
{
	const bool with_pmin = (argc == 10 && std::string(argv[7]) == "-p");
	if (argc != 7 && argc != 8 && !with_pmin)
	{
		std::cerr << "usage: ctsieve_ref m b_min b_max p_max_pos p_max_neg out.txt [dir | -p p_min_pos p_min_neg]\n";
		return 1;
	}
	const int m = std::atoi(argv[1]);
	const uint64_t b_min = std::stoull(argv[2]), b_max = std::stoull(argv[3]);
	const uint64_t p_max_pos = std::stoull(argv[4]), p_max_neg = std::stoull(argv[5]);
	Sieve s(m, b_min, b_max);
	s._p_min_pos = 3 * (1ull << m) + 1; s._p_min_neg = 11;     // the ini start (ctsieve.cpp:797)
	if (with_pmin) { s._p_min_pos = std::stoull(argv[8]); s._p_min_neg = std::stoull(argv[9]); }
	try
	{
		s.check_pos(p_max_pos);
		s.check_neg(p_max_neg);
	}
	catch (const std::runtime_error &e) { std::cerr << e.what() << std::endl; return 2; }
	std::ofstream out(argv[6]);
	for (size_t i = 0; i < s._bsieve.size(); ++i) if (!s._bsieve[i]) out << b_min + i << "\n";
	out.close();
	if (argc == 8)
	{
		if (chdir(argv[7]) != 0) { std::cerr << "cannot enter " << argv[7] << std::endl; return 3; }
		s.write(true);
	}
	return 0;
}
