// Pass kernel variants, part 0 of the table pass_variants() assembles
// (gl_ntt_kernels.hip): powers of two, up to 64 points.  One translation unit
// per part, so the instantiations compile in parallel.
#include "gl_pass.hpp"

namespace gl {

static const PassVariant kPart[] = {
    GL_VARIANT(2, 1),
    GL_VARIANT(4, 1),
    GL_VARIANT(8, 1),
    GL_VARIANT(16, 1),
    GL_VARIANT(32, 1),
    GL_VARIANT(64, 1),
    GL_VARIANT(2, 2),
    GL_VARIANT(4, 2),
    GL_VARIANT(4, 4),
    GL_VARIANT(8, 4),
};

const PassVariant *pass_variants_part0(int *count) {
    *count = (int)(sizeof(kPart) / sizeof(kPart[0]));
    return kPart;
}

}  // namespace gl
