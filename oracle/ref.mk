# ref.mk — builds, from the reference's own sources where they lie (REF, default /root/reference),
# two checkers into oracle/_ref/ (git-ignored):
#  - libsimple_ref.so: the reference's numerical CPU sources plus oracle/ref_driver.cpp (C entry
#    points), used by the CPU tests to check the oracle against the reference itself;
#  - adapter_check: the reference-side adapter simple-hpc_amd/integration/solver_sg.cpp (the
#    reference's HipSolver over this repository's C ABI), compiled against the reference's
#    headers and linked with libsimple_ref.so and build/libsimple_gpu.so, plus the per-phase and
#    whole-run driver oracle/adapter_check.cpp; a GPU test runs it.
# Only the numerical sources are compiled (no main.cpp, no io.cpp, no MPI, no CMake, no fetched
# test framework).
REF ?= /root/reference
HERE := $(dir $(lastword $(MAKEFILE_LIST)))
OUT ?= $(HERE)_ref
ROOT := $(abspath $(HERE)..)
CXX ?= g++
CXXFLAGS ?= -std=c++17 -O2 -fopenmp -ffp-contract=off -fPIC
SRCS := coefficients.cpp momentum_u.cpp momentum_v.cpp pressure.cpp multigrid.cpp solver.cpp \
        decomp.cpp params.cpp config.cpp

all: $(OUT)/libsimple_ref.so $(OUT)/adapter_check

$(OUT)/libsimple_ref.so: $(addprefix $(REF)/src/,$(SRCS)) $(HERE)ref_driver.cpp
	mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -shared -I$(REF)/include -o $@ $^

$(OUT)/adapter_check: $(HERE)adapter_check.cpp $(ROOT)/simple-hpc_amd/integration/solver_sg.cpp \
                      $(OUT)/libsimple_ref.so $(ROOT)/build/libsimple_gpu.so $(ROOT)/include/simple_gpu.h
	$(CXX) $(CXXFLAGS) -I$(REF)/include -I$(ROOT)/include -o $@ $(HERE)adapter_check.cpp \
	    $(ROOT)/simple-hpc_amd/integration/solver_sg.cpp -L$(OUT) -lsimple_ref \
	    -L$(ROOT)/build -lsimple_gpu -Wl,-rpath,'$$ORIGIN' -Wl,-rpath,'$$ORIGIN/../../build'

.PHONY: all
