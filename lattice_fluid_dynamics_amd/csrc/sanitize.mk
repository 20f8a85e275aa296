# Host code under AddressSanitizer + UndefinedBehaviorSanitizer (CPU only).
#
#   make -C lattice_fluid_dynamics_amd/csrc -f sanitize.mk check
#
# builds the C host library (lbm.c, lbm_batch.c, lbm_monitor.c, lbm_model.c,
# warmstart.c, ltws.c) and the oracle with gcc's ASan/UBSan, links the host
# objects with the unchanged HIP kernel objects of the normal build into
# ../liblbm_amd_san.so, and runs the CPU test suite against both sanitized
# libraries (LBM_AMD_LIBRARY / LBM_ORACLE_LIBRARY, runtimes preloaded into
# python).  Device code is not instrumented; the sanitizer flags go only on
# gcc lines.  Any report stops the run (halt_on_error); leak checking is off
# because the interpreter itself does not free everything at exit.
CC := gcc
SANFLAGS := -fsanitize=address,undefined -fno-sanitize-recover=undefined -fno-omit-frame-pointer -g
CFLAGS := -O1 -fPIC -std=c11 -ffp-contract=off -Wall -Wno-unused-function -Wno-unknown-pragmas $(SANFLAGS)
SANB := _build/san
HOST := lbm lbm_monitor lbm_batch lbm_model warmstart ltws
HOSTOBJ := $(HOST:%=$(SANB)/%.o)
HIPOBJ := _build/lbm_kernels.o _build/warmstart_kernels.o _build/mesh_kernels.o _build/superres.o
OUT := ../liblbm_amd_san.so
ORACLE := ../../oracle
ORACLE_OUT := $(ORACLE)/liblbm_oracle_san.so
ROCM ?= /opt/rocm

all: $(OUT) $(ORACLE_OUT)

$(SANB):
	mkdir -p $(SANB)

$(SANB)/%.o: %.c $(wildcard *.h) ../../include/lbm.h | $(SANB)
	$(CC) $(CFLAGS) -c $< -o $@

# the kernel objects come from the normal build
$(HIPOBJ):
	$(MAKE) -f Makefile

$(OUT): $(HOSTOBJ) $(HIPOBJ)
	$(CC) -shared $(SANFLAGS) -o $@ $(HOSTOBJ) $(HIPOBJ) -L$(ROCM)/lib -Wl,-rpath,$(ROCM)/lib -lamdhip64 -lstdc++ -lm

$(ORACLE_OUT): $(ORACLE)/oracle_geom.c $(ORACLE)/oracle_particles.c $(ORACLE)/oracle_dyn.c $(ORACLE)/oracle.h | $(SANB)
	$(CC) $(CFLAGS) -fopenmp -c $(ORACLE)/oracle_geom.c -o $(SANB)/or_geom.o
	$(CC) $(CFLAGS) -fopenmp -c $(ORACLE)/oracle_particles.c -o $(SANB)/or_part.o
	$(CC) $(CFLAGS) -fopenmp -c $(ORACLE)/oracle_dyn.c -o $(SANB)/or_dyn32.o
	$(CC) $(CFLAGS) -fopenmp -DREAL_IS_DOUBLE -c $(ORACLE)/oracle_dyn.c -o $(SANB)/or_dyn64.o
	$(CC) -shared $(SANFLAGS) -fopenmp -o $@ $(SANB)/or_geom.o $(SANB)/or_part.o $(SANB)/or_dyn32.o $(SANB)/or_dyn64.o -lm

check: $(OUT) $(ORACLE_OUT)
	cd ../.. && LBM_AMD_LIBRARY=$(abspath $(OUT)) LBM_ORACLE_LIBRARY=$(abspath $(ORACLE_OUT)) \
	  ASAN_OPTIONS=detect_leaks=0:halt_on_error=1:abort_on_error=1 \
	  UBSAN_OPTIONS=halt_on_error=1:print_stacktrace=1 \
	  LD_PRELOAD="$$($(CC) -print-file-name=libasan.so) $$($(CC) -print-file-name=libubsan.so)" \
	  python3 -m pytest tests -m "not gpu" -q -p no:cacheprovider

clean:
	rm -rf $(SANB) $(OUT) $(ORACLE_OUT)

.PHONY: all check clean
