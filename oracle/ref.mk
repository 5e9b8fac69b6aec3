# Builds, from the reference's own sources compiled where they lie under
# /root/reference (test infrastructure only; run in the build container, the
# GPU box has no reference tree and uses the prebuilt files):
#   _ref/libreflist.so  its list producer + oracle/ref_harness.c (captures the
#                       leaf pairs handed to the device library)
#   _ref/link_host      its fmm_prepare + fmm_task + oracle/link_harness.c,
#                       linked against the drop-in libp2p_hip.so (needs it built)
# Flags: the default Makefile's INTXYZ + LONGSHORT + GPUP2P, without
# MTHKDTree (its threaded build numbers leaves nondeterministically).
REF      ?= /root/reference
OUT      := _ref
CC       ?= gcc
CXX      ?= g++
DEFS     := -DINTXYZ -DLONGSHORT -DGPUP2P
CFLAGS   := -O2 -fPIC -fopenmp -ffp-contract=off -fno-fast-math -w $(DEFS) -I$(REF)/inc
SRCS_C   := fmm.c operator.c memory.c tasklog.c
REFOBJS  := $(addprefix $(OUT)/,$(SRCS_C:.c=.o)) $(OUT)/utility.o
OBJS     := $(REFOBJS) $(OUT)/ref_harness.o
P2PLIB   := ../photons-on-mt3000_amd/lib

all: $(OUT)/libreflist.so $(OUT)/link_host

$(OUT):
	mkdir -p $(OUT)
$(OUT)/%.o: $(REF)/src/%.c | $(OUT)
	$(CC) $(CFLAGS) -c $< -o $@
$(OUT)/utility.o: $(REF)/src/utility.cc | $(OUT)
	$(CXX) $(CFLAGS) -c $< -o $@
$(OUT)/ref_harness.o: ref_harness.c | $(OUT)
	$(CC) $(CFLAGS) -std=gnu11 -c $< -o $@
$(OUT)/libreflist.so: $(OBJS)
	$(CXX) -shared -fopenmp -o $@ $(OBJS) -lm
$(OUT)/link_harness.o: link_harness.c | $(OUT)
	$(CC) $(CFLAGS) -std=gnu11 -c $< -o $@
$(OUT)/link_host: $(REFOBJS) $(OUT)/link_harness.o $(P2PLIB)/libp2p_hip.so
	$(CXX) -fopenmp -o $@ $(REFOBJS) $(OUT)/link_harness.o -L$(P2PLIB) -lp2p_hip \
	    -Wl,-rpath,'$$ORIGIN/../$(P2PLIB)' -lm

clean:
	rm -rf $(OUT)
.PHONY: all clean
