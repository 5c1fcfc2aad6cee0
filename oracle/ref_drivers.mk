# ref_drivers.mk -- builds the reference's own FD and SEM drivers, unchanged,
# against this library's drop-in headers (proxy-geos-hc_amd/host) and
# libphc.so, into oracle/_ref/.  This checks that the reference's callers
# (src/fd/FDTDmain.cpp; src/sem/main/SEMproxy.cpp and main.cpp) compile and
# run on the drop-in classes with no change.  The SEM files are compiled from
# standard input, so that their #include "SEMproxy.hpp" resolves to the
# drop-in header rather than to the reference's header beside them.
#
#   make -f oracle/ref_drivers.mk REF=/root/reference
REF ?= /root/reference
ROOT := $(abspath $(dir $(lastword $(MAKEFILE_LIST)))/..)
OUT := $(ROOT)/oracle/_ref
CXXFLAGS := -O2 -std=c++17 -I$(ROOT)/include -I$(ROOT)/proxy-geos-hc_amd/host
LDFLAGS := -L$(ROOT)/proxy-geos-hc_amd -lphc -Wl,-rpath,'$$ORIGIN/../../proxy-geos-hc_amd'

all: $(OUT)/fd_ref_driver $(OUT)/sem_ref_driver

$(OUT):
	mkdir -p $(OUT)

$(OUT)/fd_ref_driver: $(REF)/src/fd/FDTDmain.cpp | $(OUT)
	g++ $(CXXFLAGS) $< -o $@ $(LDFLAGS)

$(OUT)/sem_ref_driver: $(REF)/src/sem/main/SEMproxy.cpp $(REF)/src/sem/main/main.cpp | $(OUT)
	g++ $(CXXFLAGS) -c -x c++ - -o $(OUT)/SEMproxy.o < $(REF)/src/sem/main/SEMproxy.cpp
	g++ $(CXXFLAGS) -c -x c++ - -o $(OUT)/main.o < $(REF)/src/sem/main/main.cpp
	g++ $(OUT)/SEMproxy.o $(OUT)/main.o -o $@ $(LDFLAGS)
	rm -f $(OUT)/SEMproxy.o $(OUT)/main.o

.PHONY: all
