# sanitize.mk -- sanitizer builds of the oracle, run on the CPU (test
# infrastructure; outputs into oracle/_san/, which is not committed).
#   asan: gcc -fsanitize=address,undefined, every oracle loop on 4 threads
#   tsan: clang -fsanitize=thread with the LLVM OpenMP runtime and its race
#         annotations (libarcher), so that OpenMP barriers and reductions are
#         understood and only true races are reported
#
#   make -f oracle/sanitize.mk asan tsan tsan-planted
ROOT := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))/..)
OUT := $(ROOT)/oracle/_san
SRC := $(ROOT)/oracle/sanitize_check.c $(ROOT)/oracle/fd_oracle.c $(ROOT)/oracle/sem_oracle.c $(ROOT)/oracle/sem2d_oracle.c
LLVM ?= /opt/rocm/llvm

all: asan tsan tsan-planted

$(OUT):
	mkdir -p $(OUT)

$(OUT)/check_asan: $(SRC) $(ROOT)/oracle/oracle.h | $(OUT)
	gcc -O1 -g -fopenmp -ffp-contract=off -fsanitize=address,undefined -fno-sanitize-recover=all \
	    -fno-omit-frame-pointer -I$(ROOT)/oracle $(SRC) -o $@ -lm

$(OUT)/check_tsan: $(SRC) $(ROOT)/oracle/oracle.h | $(OUT)
	$(LLVM)/bin/clang -O1 -g -fopenmp -ffp-contract=off -fsanitize=thread -I$(ROOT)/oracle $(SRC) -o $@ -lm \
	    -Wl,-rpath,$(LLVM)/lib

# the same with a race planted in the colour-parallel element loop: the run
# must fail (this target succeeds when ThreadSanitizer reports the race)
$(OUT)/check_tsan_planted: $(SRC) $(ROOT)/oracle/oracle.h | $(OUT)
	$(LLVM)/bin/clang -O1 -g -fopenmp -ffp-contract=off -fsanitize=thread -DORC_PLANT_SCATTER_RACE \
	    -I$(ROOT)/oracle $(SRC) -o $@ -lm -Wl,-rpath,$(LLVM)/lib

tsan-planted: $(OUT)/check_tsan_planted
	! OMP_NUM_THREADS=4 OMP_TOOL_LIBRARIES=$(LLVM)/lib/libarcher.so \
	    TSAN_OPTIONS="halt_on_error=1 ignore_noninstrumented_modules=1" $(OUT)/check_tsan_planted

asan: $(OUT)/check_asan
	OMP_NUM_THREADS=4 ASAN_OPTIONS=detect_leaks=1 $(OUT)/check_asan

tsan: $(OUT)/check_tsan
	OMP_NUM_THREADS=4 OMP_TOOL_LIBRARIES=$(LLVM)/lib/libarcher.so \
	    TSAN_OPTIONS="halt_on_error=1 ignore_noninstrumented_modules=1" $(OUT)/check_tsan

.PHONY: all asan tsan tsan-planted
