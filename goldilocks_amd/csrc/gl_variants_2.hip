// Pass kernel variants, part 2 of the table pass_variants() assembles
// (gl_ntt_kernels.hip): powers of two, 1024 to 4096 points.  One translation
// unit per part, so the instantiations compile in parallel.
#include "gl_pass.hpp"

namespace gl {

static const PassVariant kPart[] = {
    GL_VARIANT(32, 32),
    GL_VARIANT(64, 32),
    GL_VARIANT(64, 64),
};

const PassVariant *pass_variants_part2(int *count) {
    *count = (int)(sizeof(kPart) / sizeof(kPart[0]));
    return kPart;
}

}  // namespace gl
