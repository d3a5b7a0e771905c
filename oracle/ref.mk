# Builds the reference's bundled CPU presolver, PSLP 0.0.11
# (third_party/pslp, Apache-2.0), from its sources where they lie, with the
# reference's flags (solver_interface/CMakeLists.txt:22-36), together with
# pslp_driver.c.  CPU baseline only: output goes to _ref/ (not committed).
REF ?= /root/reference
PSLP := $(REF)/third_party/pslp
PSLP_SRC := $(filter-out $(PSLP)/src/core/Debugger.c,$(wildcard $(PSLP)/src/core/*.c $(PSLP)/src/explorers/*.c))
PSLP_INC := -I$(PSLP)/include/PSLP -I$(PSLP)/include/core -I$(PSLP)/include/data_structures -I$(PSLP)/include/explorers
CFLAGS := -O3 -std=c99 -fPIC -DNDEBUG -D_POSIX_C_SOURCE=200809L -DPSLP_VERSION=\"0.0.11\" $(PSLP_INC)

_ref/libpslp.so: pslp_driver.c $(PSLP_SRC)
	mkdir -p _ref
	$(CC) $(CFLAGS) -shared -o $@ pslp_driver.c $(PSLP_SRC) -lpthread -lm

.PHONY: clean
clean:
	rm -rf _ref
