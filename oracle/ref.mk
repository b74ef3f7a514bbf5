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

all: $(OUT)/runnable $(OUT)/refprobe $(OUT)/refprobe64

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

.PHONY: all
