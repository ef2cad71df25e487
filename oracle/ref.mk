# TEST INFRASTRUCTURE ONLY: the reference's numeric sources (Ruiz, the supernodal kernel, the row kernel, the sweeps and the
# solve plans)
# compiled where they lie under /root/reference into _ref/libfplus_ref.so, to validate the oracle.  They need only libm:
# without a CBLAS the supernodal kernel is the column-at-a-time loop form.  The reference's flags (CMakeLists.txt:24-31,
# :98-101), -march left at the default so the library runs on any x86-64 host; -std=c99 does not contract.
REF     ?= /root/reference
CC      ?= gcc
HERE    := $(dir $(abspath $(lastword $(MAKEFILE_LIST))))
OUT     := $(HERE)_ref
CFLAGS  := -std=c99 -O3 -fno-math-errno -fno-signed-zeros -DNDEBUG -fPIC -I$(REF)/include -I$(REF)/src

all: $(OUT)/libfplus_ref.so $(OUT)/libfplus_ref_row_omp.so

$(OUT)/libfplus_ref.so: $(REF)/src/equil.c $(REF)/src/factor_sn.c $(REF)/src/factor_row.c $(REF)/src/solve.c $(REF)/src/solve_plan.c
	mkdir -p $(OUT)
	$(CC) $(CFLAGS) -c $(REF)/src/equil.c -o $(OUT)/equil.o
	$(CC) $(CFLAGS) -c $(REF)/src/factor_sn.c -o $(OUT)/factor_sn.o
	$(CC) $(CFLAGS) -c $(REF)/src/factor_row.c -o $(OUT)/factor_row.o
	$(CC) $(CFLAGS) -ffp-contract=off -c $(REF)/src/solve.c -o $(OUT)/solve.o
	$(CC) $(CFLAGS) -ffp-contract=off -c $(REF)/src/solve_plan.c -o $(OUT)/solve_plan.o
	$(CC) -shared -o $@ $(OUT)/equil.o $(OUT)/factor_sn.o $(OUT)/factor_row.o $(OUT)/solve.o $(OUT)/solve_plan.o -lm
	rm -f $(OUT)/*.o

# the row kernel as the reference ships it for timing: OpenMP on (its default, CMakeLists.txt:5), vectorised for any
# x86-64 host with AVX2 (the reference's -march=native would tie the library to this container's CPU); the CPU baseline
# of the row-kernel benchmark
$(OUT)/libfplus_ref_row_omp.so: $(REF)/src/factor_row.c
	mkdir -p $(OUT)
	$(CC) $(CFLAGS) -march=x86-64-v3 -fopenmp -shared -o $@ $(REF)/src/factor_row.c -lm

.PHONY: all
