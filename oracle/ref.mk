# Builds from the reference's sources where they lie, into _ref/ (not
# committed):
#  * its bundled CPU presolver, PSLP 0.0.11 (third_party/pslp, Apache-2.0),
#    with the reference's flags (solver_interface/CMakeLists.txt:22-36),
#    together with pslp_driver.c -- the CPU baseline;
#  * its MPS reader (cpp/mps_reader: three sources, no dependencies; built
#    there by cpp/mps_reader/CMakeLists.txt with C++17 and default flags)
#    together with mps_ref_driver.cpp -- the check of the MPS restatement.
REF ?= /root/reference
CXX ?= g++
PSLP := $(REF)/third_party/pslp
PSLP_SRC := $(filter-out $(PSLP)/src/core/Debugger.c,$(wildcard $(PSLP)/src/core/*.c $(PSLP)/src/explorers/*.c))
PSLP_INC := -I$(PSLP)/include/PSLP -I$(PSLP)/include/core -I$(PSLP)/include/data_structures -I$(PSLP)/include/explorers
CFLAGS := -O3 -std=c99 -fPIC -DNDEBUG -D_POSIX_C_SOURCE=200809L -DPSLP_VERSION=\"0.0.11\" $(PSLP_INC)

MPS := $(REF)/cpp/mps_reader
MPS_SRC := $(MPS)/src/io/mpsreader.cpp $(MPS)/src/core/lp_model.cpp $(MPS)/src/core/sparse_matrix.cpp

all: _ref/libpslp.so _ref/libmpsref.so

_ref/libmpsref.so: mps_ref_driver.cpp $(MPS_SRC)
	mkdir -p _ref
	$(CXX) -O2 -std=c++17 -fPIC -shared -I$(MPS)/include -o $@ mps_ref_driver.cpp $(MPS_SRC)

_ref/libpslp.so: pslp_driver.c $(PSLP_SRC)
	mkdir -p _ref
	$(CC) $(CFLAGS) -shared -o $@ pslp_driver.c $(PSLP_SRC) -lpthread -lm

.PHONY: all clean
clean:
	rm -rf _ref
