// Pass kernel variants, part 1 of the table pass_variants() assembles
// (gl_ntt_kernels.hip): powers of two, 64 to 512 points.  One translation unit
// per part, so the instantiations compile in parallel.
#include "gl_pass.hpp"

namespace gl {

static const PassVariant kPart[] = {
    GL_VARIANT(8, 8),
    GL_VARIANT(16, 8),
    GL_VARIANT(16, 16),
    GL_VARIANT(32, 16),
};

const PassVariant *pass_variants_part1(int *count) {
    *count = (int)(sizeof(kPart) / sizeof(kPart[0]));
    return kPart;
}

}  // namespace gl
