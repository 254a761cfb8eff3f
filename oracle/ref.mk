# TEST INFRASTRUCTURE ONLY.
# Compiles the reference engine (sblanas/pythia) from its own sources where
# they lie under $(REF) into oracle/_ref/ -- nothing is copied into the
# repository.  HDF5 is off (its tarball is absent from the reference); the
# vendored libconfig 1.2 and bzip2 1.0.5 are compiled from $(REF)/externals.
# The driver oracle/ref_driver.cpp (ours) runs the reference's HashJoinOp /
# AggregateSum / AggregateCount / MergeOp on generated inputs.
REF ?= /root/reference
OUT := _ref
CXX ?= g++
CC := gcc
INC := -I$(REF) -I$(REF)/externals/libconfig-1.2 -I$(REF)/externals/bzip2-1.0.5
CXXFLAGS := -std=gnu++98 -O3 -m64 -w -DSTATS_ALLOCATE -pthread $(INC)
CFLAGS := -O2 -w $(INC)

PY_SRC := schema.cpp hash.cpp ProcessorMap.cpp Barrier.cpp perfcounters.cpp query.cpp \
	util/hashtable.cpp util/buffer.cpp visitors/recursivedestroy.cpp visitors/recursivefree.cpp \
	visitors/prettyprint.cpp visitors/threadinit.cpp visitors/threadclose.cpp operators/base.cpp \
	operators/mapwrapper.cpp operators/filter.cpp operators/sortlimit.cpp operators/genericaggregate.cpp \
	operators/aggregatecount.cpp operators/aggregatesum.cpp operators/scan.cpp operators/partitionedscan.cpp \
	operators/parallelscan.cpp operators/merge.cpp operators/join.cpp operators/shuffle.cpp \
	operators/cycleaccountant.cpp util/affinitizer.cpp operators/project.cpp comparator.cpp \
	conjunctionevaluator.cpp rawcompfns.cpp operators/memsegmentwriter.cpp operators/loaders/table.cpp \
	operators/loaders/parser.cpp operators/loaders/loader.cpp operators/checker_callstate.cpp \
	operators/printer_tuplecount.cpp operators/generator_int.cpp operators/printer_perfcnt.cpp \
	operators/bitentropy.cpp operators/threadidprepend.cpp util/numaasserts.cpp util/numaallocate.cpp \
	operators/consume.cpp operators/sortandrangepartition.cpp operators/partition.cpp operators/indexjoin.cpp
PY_CSRC := operators/loaders/sfmt/SFMT.c
LC_SRC := libconfig.c grammar.c scanner.c
BZ_SRC := blocksort.c huffman.c crctable.c randtable.c compress.c decompress.c bzlib.c

flat = $(subst /,__,$(1))
PY_OBJ := $(foreach s,$(PY_SRC),$(OUT)/obj/$(call flat,$(s)).o) $(OUT)/obj/$(call flat,$(PY_CSRC)).o
LC_OBJ := $(foreach s,$(LC_SRC),$(OUT)/obj/libconfig__$(s).o) $(OUT)/obj/libconfig__libconfigcpp.o
BZ_OBJ := $(foreach s,$(BZ_SRC),$(OUT)/obj/bzip2__$(s).o)

all: $(OUT)/refjoin $(OUT)/refgpujoin

$(OUT)/obj:
	mkdir -p $@

define py_rule
$(OUT)/obj/$(call flat,$(1)).o: $(REF)/$(1) | $(OUT)/obj
	$(CXX) $(CXXFLAGS) -c $$< -o $$@
endef
$(foreach s,$(PY_SRC),$(eval $(call py_rule,$(s))))

# The reference builds C sources with CC=$(CXX) (Makefile:50-51).
$(OUT)/obj/$(call flat,$(PY_CSRC)).o: $(REF)/$(PY_CSRC) | $(OUT)/obj
	$(CXX) -x c++ $(CXXFLAGS) -c $< -o $@

$(OUT)/obj/libconfig__%.c.o: $(REF)/externals/libconfig-1.2/%.c | $(OUT)/obj
	$(CC) $(CFLAGS) -c $< -o $@
$(OUT)/obj/libconfig__libconfigcpp.o: $(REF)/externals/libconfig-1.2/libconfigcpp.c++ | $(OUT)/obj
	$(CXX) -x c++ $(CXXFLAGS) -c $< -o $@
$(OUT)/obj/bzip2__%.c.o: $(REF)/externals/bzip2-1.0.5/%.c | $(OUT)/obj
	$(CC) $(CFLAGS) -D_FILE_OFFSET_BITS=64 -c $< -o $@

$(OUT)/refjoin: ref_driver.cpp ref_memscan.h ../include/pgpu_gen.h $(PY_OBJ) $(LC_OBJ) $(BZ_OBJ)
	$(CXX) -std=gnu++98 -O3 -w -DSTATS_ALLOCATE -pthread $(INC) -I../include ref_driver.cpp $(PY_OBJ) $(LC_OBJ) $(BZ_OBJ) -lrt -o $@

# The reference engine with the GPU operators' binding (integration/) linked
# in, against libpgpu.so (built first by the package's build step).
$(OUT)/refgpujoin: ref_gpu_driver.cpp ref_memscan.h ../integration/gpujoin.cpp ../integration/gpujoin.h ../include/pgpu.h ../pythia_amd/libpgpu.so $(PY_OBJ) $(LC_OBJ) $(BZ_OBJ)
	$(CXX) -std=gnu++98 -O2 -w -DSTATS_ALLOCATE -pthread $(INC) -I../include ref_gpu_driver.cpp ../integration/gpujoin.cpp $(PY_OBJ) $(LC_OBJ) $(BZ_OBJ) -L../pythia_amd -lpgpu -Wl,-rpath,'$$ORIGIN/../../pythia_amd' -lrt -o $@

clean:
	rm -rf $(OUT)

.PHONY: all clean
