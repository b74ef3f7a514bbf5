# Compiles the reference CPU path from its own sources where they lie under
# REF (default /root/reference), with the MPICH compiler wrappers shipped in
# the image.  Outputs go to oracle/_ref/ only (git-ignored; they travel to the
# GPU box next to the product .so).  Nothing from REF is copied.
#
#   oracle/_ref/runnable   the reference benchmark program, VERTEX_REORDERING=2,
#                          64 roots (REAL_BENCHMARK), quiet build
#   oracle/_ref/refprobe   our dump driver (oracle/ref_probe.cc) linked against
#                          the reference generator/constructor/BFS/validator,
#                          VERBOSE build so per-level directions are printed
#   oracle/_ref/runnable_gpu
#                          the reference driver mpi/main.cc with g5::BfsOnGPU
#                          (include/g5_bfs.hpp over graph500_amd/lib/libg5.so)
#                          in place of BfsOnCPU: sed rewrites the class name at
#                          its three uses on the way into the compiler (no
#                          edited copy is written anywhere); everything else --
#                          generation, find_roots, the timing, get_pred, the
#                          reference's own validate_bfs_result and result
#                          block -- is the reference's code, 64 roots, VR=2
REF ?= /root/reference
MPIBIN ?= /opt/conda/bin
HERE := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT := $(HERE)_ref
export MPICH_CXX := g++
export MPICH_CC := gcc
MPICXX := $(MPIBIN)/mpicxx
MPICC := $(MPIBIN)/mpicc

BASE := -fopenmp -g -Drestrict=__restrict__ -D__STDC_CONSTANT_MACROS -D__STDC_LIMIT_MACROS \
        -D__STDC_FORMAT_MACROS -ffast-math -msse4.2 -O3 -DNDEBUG -w
RUNNABLE_FLAGS := $(BASE) -DVERVOSE_MODE=0 -DVERTEX_REORDERING=2 -DREAL_BENCHMARK
PROBE_FLAGS := $(BASE) -DVERVOSE_MODE=1 -DVERTEX_REORDERING=2

INCLUDE := $(abspath $(HERE)../include)
LIBDIR := $(abspath $(HERE)../graph500_amd/lib)

all: $(OUT)/runnable $(OUT)/refprobe $(OUT)/refprobe64 $(OUT)/runnable_gpu

$(OUT)/.dir:
	mkdir -p $(OUT) && touch $@

$(OUT)/splittable_mrg.o: $(REF)/generator/splittable_mrg.c $(OUT)/.dir
	$(MPICC) -c $(BASE) $< -o $@

$(OUT)/low_level_func.o: $(REF)/mpi/low_level_func.cc $(OUT)/.dir
	$(MPICXX) -c $(BASE) $< -o $@

$(OUT)/main.o: $(REF)/mpi/main.cc $(OUT)/.dir
	$(MPICXX) -c $(RUNNABLE_FLAGS) $< -o $@

$(OUT)/runnable: $(OUT)/main.o $(OUT)/low_level_func.o $(OUT)/splittable_mrg.o
	$(MPICXX) -fopenmp -o $@ $^ -lm -lnuma

$(OUT)/ref_probe.o: $(HERE)ref_probe.cc $(OUT)/.dir
	$(MPICXX) -c $(PROBE_FLAGS) -I$(REF)/mpi $< -o $@

$(OUT)/ref_probe64.o: $(HERE)ref_probe.cc $(OUT)/.dir
	$(MPICXX) -c $(PROBE_FLAGS) -DREAL_BENCHMARK -I$(REF)/mpi $< -o $@

$(OUT)/refprobe: $(OUT)/ref_probe.o $(OUT)/low_level_func.o $(OUT)/splittable_mrg.o
	$(MPICXX) -fopenmp -o $@ $^ -lm -lnuma

$(OUT)/refprobe64: $(OUT)/ref_probe64.o $(OUT)/low_level_func.o $(OUT)/splittable_mrg.o
	$(MPICXX) -fopenmp -o $@ $^ -lm -lnuma

# The drop-in swap: BfsOnCPU -> g5::BfsOnGPU (mpi/main.cc:49, :58).
SWAP := -e 's/BfsOnCPU::printInformation/g5::BfsOnGPU::printInformation/' \
        -e 's/BfsOnCPU\* benchmark = new BfsOnCPU()/g5::BfsOnGPU* benchmark = new g5::BfsOnGPU()/'

$(OUT)/main_gpu.o: $(REF)/mpi/main.cc $(INCLUDE)/g5_bfs.hpp $(INCLUDE)/g5.h $(OUT)/.dir
	sed $(SWAP) $< | grep -c 'g5::BfsOnGPU' | grep -qx 2
	sed $(SWAP) $< | $(MPICXX) -x c++ -c $(RUNNABLE_FLAGS) -I$(REF)/mpi -I$(INCLUDE) -include g5_bfs.hpp - -o $@

# Linked by g++ itself with the MPI libraries named: the MPI wrapper would put
# its own (older) libstdc++ first, which the ROCm runtime under libg5.so
# cannot use; the system one comes first in the run path.
MPILIB := $(abspath $(MPIBIN)/../lib)
$(OUT)/runnable_gpu: $(OUT)/main_gpu.o $(OUT)/low_level_func.o $(OUT)/splittable_mrg.o $(LIBDIR)/libg5.so
	g++ -fopenmp -o $@ $(OUT)/main_gpu.o $(OUT)/low_level_func.o $(OUT)/splittable_mrg.o \
	  -L$(LIBDIR) -lg5 $(MPILIB)/libmpicxx.so $(MPILIB)/libmpi.so \
	  -Wl,-rpath,'$$ORIGIN/../../graph500_amd/lib:/usr/lib/x86_64-linux-gnu:$(MPILIB)' \
	  -Wl,-rpath-link,/opt/rocm/lib -lm -lnuma

.PHONY: all
