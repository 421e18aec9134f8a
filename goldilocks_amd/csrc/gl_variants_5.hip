// Pass kernel variants, part 5 of the table pass_variants() assembles
// (gl_ntt_kernels.hip): odd factors after 8, 4 and 2.  One translation unit
// per part, so the instantiations compile in parallel.
#include "gl_pass.hpp"

namespace gl {

static const PassVariant kPart[] = {
    GL_VARIANT(8, 3),
    GL_VARIANT(8, 5),
    GL_VARIANT(8, 15),
    GL_VARIANT(8, 17),
    GL_VARIANT(4, 3),
    GL_VARIANT(4, 5),
    GL_VARIANT(4, 15),
    GL_VARIANT(4, 17),
    GL_VARIANT(2, 3),
    GL_VARIANT(2, 5),
    GL_VARIANT(2, 15),
    GL_VARIANT(2, 17),
};

const PassVariant *pass_variants_part5(int *count) {
    *count = (int)(sizeof(kPart) / sizeof(kPart[0]));
    return kPart;
}

}  // namespace gl
