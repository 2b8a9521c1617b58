# Compiles the reference's own rolling hashes, unmodified where they lie in
# the reference tree, behind oracle/ref_hash_driver.cpp into
# oracle/_ref/libref_hashes.so (test infrastructure: the CPU tests compare the
# oracle's window hashes with them).  Nothing else of the reference is built:
# ffbf.cpp does not compile with this toolchain (ffbf.cpp:501 compares the
# `read` function pointer) and greptest.cu is CUDA.
#
# -O0: for some window lengths the headers shift a 32-bit word by 32 or more
# (rol32(x, LEN) with LEN >= 32, rol32(x, 2*LEN % 32) when that is 0) and a
# 64-bit word by 64 (rol64(x, 64)), which C++ leaves undefined.  Unoptimised
# x86-64 code executes those shifts with the count taken mod 32 (mod 64),
# which makes rol32/rol64 rotations by the amount mod 32 (mod 64) -- the
# semantics the oracle restates; an optimising build may fold them otherwise.
#
#   make -f ref.mk            (from oracle/; REF= overrides the reference path)
REF ?= /root/reference
CXX ?= g++

_ref/libref_hashes.so: ref_hash_driver.cpp
	mkdir -p _ref
	$(CXX) -O0 -g -std=c++14 -fPIC -shared -Wall -I$(REF)/src/hashes -o $@ ref_hash_driver.cpp

clean:
	rm -rf _ref

.PHONY: clean
