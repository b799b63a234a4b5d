# Builds the reference PCAone CLI from its own sources under /root/reference into
# oracle/_ref/ (git-ignored, never committed).  This recipe is ours: it does not
# call the reference's Makefile.  The binary is a CPU checker and the CPU
# baseline timed beside the GPU path; nothing in the product links to it.
#
#   make -f oracle/ref.mk -j8
#
# Needs only what this image has: g++ (C++17, OpenMP), gcc, zlib.  zstd and bgen
# are compiled from the reference's vendored sources.
REF      ?= /root/reference
OUT      ?= oracle/_ref
OBJDIR   := $(OUT)/obj
CXX      := g++
CC       := gcc
# x86-64-v3 (AVX2+FMA) instead of -march=native so the binary runs on the
# GPU host's CPU as well as this one.
ARCHFLAG := -march=x86-64-v3
CXXFLAGS := -O3 -std=c++17 -DNDEBUG -m64 $(ARCHFLAG) -fopenmp -DVERSION=\"0.6.0\" -w
CFLAGS   := -O3 -DNDEBUG $(ARCHFLAG) -DZSTD_DISABLE_ASM -DZSTD_MULTITHREAD=0 -w
INC      := -I$(REF)/external -I$(REF)/external/zstd/lib -I$(REF)/external/zstd/lib/common

PCA_SRC  := $(wildcard $(REF)/src/*.cpp)
BGEN_SRC := $(wildcard $(REF)/external/bgen/*.cpp)
ZSTD_SRC := $(wildcard $(REF)/external/zstd/lib/common/*.c) \
            $(wildcard $(REF)/external/zstd/lib/compress/*.c) \
            $(wildcard $(REF)/external/zstd/lib/decompress/*.c)

PCA_OBJ  := $(patsubst $(REF)/src/%.cpp,$(OBJDIR)/pca_%.o,$(PCA_SRC))
BGEN_OBJ := $(patsubst $(REF)/external/bgen/%.cpp,$(OBJDIR)/bgen_%.o,$(BGEN_SRC))
ZSTD_OBJ := $(patsubst $(REF)/external/zstd/lib/%.c,$(OBJDIR)/zstd_%.o,$(subst /,_,$(ZSTD_SRC)))

all: $(OUT)/PCAone

$(OBJDIR)/pca_%.o: $(REF)/src/%.cpp | $(OBJDIR)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

$(OBJDIR)/bgen_%.o: $(REF)/external/bgen/%.cpp | $(OBJDIR)
	$(CXX) $(CXXFLAGS) $(INC) -c $< -o $@

ZSTD_ALL := $(ZSTD_SRC)
define zstd_rule
$(OBJDIR)/zstd_$(subst /,_,$(patsubst $(REF)/external/zstd/lib/%,%,$(1:.c=.o))): $(1) | $(OBJDIR)
	$(CC) $(CFLAGS) $(INC) -c $$< -o $$@
endef
$(foreach s,$(ZSTD_ALL),$(eval $(call zstd_rule,$(s))))
ZSTD_OBJS := $(foreach s,$(ZSTD_ALL),$(OBJDIR)/zstd_$(subst /,_,$(patsubst $(REF)/external/zstd/lib/%,%,$(s:.c=.o))))

$(OUT)/PCAone: $(PCA_OBJ) $(BGEN_OBJ) $(ZSTD_OBJS)
	$(CXX) $(CXXFLAGS) -o $@ $^ -lz -lpthread -ldl -lm

$(OBJDIR):
	mkdir -p $@

.PHONY: all
