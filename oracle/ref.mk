# Builds oracle/_ref/libreflist.so from the reference's own list-producer
# sources, compiled where they lie under /root/reference, plus
# oracle/ref_harness.c.  Test infrastructure only; run in the build
# container (the GPU box has no reference tree and uses the prebuilt .so).
# Flags: the default Makefile's INTXYZ + LONGSHORT + GPUP2P, without
# MTHKDTree (its threaded build numbers leaves nondeterministically).
REF      ?= /root/reference
OUT      := _ref
CC       ?= gcc
CXX      ?= g++
DEFS     := -DINTXYZ -DLONGSHORT -DGPUP2P
CFLAGS   := -O2 -fPIC -fopenmp -ffp-contract=off -fno-fast-math -w $(DEFS) -I$(REF)/inc
SRCS_C   := fmm.c operator.c memory.c tasklog.c
OBJS     := $(addprefix $(OUT)/,$(SRCS_C:.c=.o)) $(OUT)/utility.o $(OUT)/ref_harness.o

all: $(OUT)/libreflist.so

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

clean:
	rm -rf $(OUT)
.PHONY: all clean
