// Pass kernel variants, part 4 of the table pass_variants() assembles
// (gl_ntt_kernels.hip): odd factors after 32 and 16.  One translation unit per
// part, so the instantiations compile in parallel.
#include "gl_pass.hpp"

namespace gl {

static const PassVariant kPart[] = {
    GL_VARIANT(32, 3),
    GL_VARIANT(32, 5),
    GL_VARIANT(32, 15),
    GL_VARIANT(32, 17),
    GL_VARIANT(16, 3),
    GL_VARIANT(16, 5),
    GL_VARIANT(16, 15),
    GL_VARIANT(16, 17),
};

const PassVariant *pass_variants_part4(int *count) {
    *count = (int)(sizeof(kPart) / sizeof(kPart[0]));
    return kPart;
}

}  // namespace gl
