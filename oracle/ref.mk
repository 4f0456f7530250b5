# ref.mk — builds oracle/_ref/libsimple_ref.so from the reference's own CPU sources where they lie
# (REF, default /root/reference), plus oracle/ref_driver.cpp. Only the numerical sources are
# compiled (no main.cpp, no io.cpp, no MPI, no CMake, no fetched test framework). The result is
# git-ignored and used only by tests to check the oracle against the reference itself.
REF ?= /root/reference
OUT ?= $(dir $(lastword $(MAKEFILE_LIST)))_ref
CXX ?= g++
CXXFLAGS ?= -std=c++17 -O2 -fopenmp -ffp-contract=off -fPIC
SRCS := coefficients.cpp momentum_u.cpp momentum_v.cpp pressure.cpp multigrid.cpp solver.cpp \
        decomp.cpp params.cpp config.cpp
HERE := $(dir $(lastword $(MAKEFILE_LIST)))

all: $(OUT)/libsimple_ref.so

$(OUT)/libsimple_ref.so: $(addprefix $(REF)/src/,$(SRCS)) $(HERE)ref_driver.cpp
	mkdir -p $(OUT)
	$(CXX) $(CXXFLAGS) -shared -I$(REF)/include -o $@ $^

.PHONY: all
