// Pass kernel variants, part 3 of the table pass_variants() assembles
// (gl_ntt_kernels.hip): odd factors 3, 5, 15, 17 alone and after 64.  One
// translation unit per part, so the instantiations compile in parallel.
#include "gl_pass.hpp"

namespace gl {

static const PassVariant kPart[] = {
    GL_VARIANT(3, 1),
    GL_VARIANT(5, 1),
    GL_VARIANT(15, 1),
    GL_VARIANT(17, 1),
    GL_VARIANT(64, 3),
    GL_VARIANT(64, 5),
    GL_VARIANT(64, 15),
    GL_VARIANT(64, 17),
};

const PassVariant *pass_variants_part3(int *count) {
    *count = (int)(sizeof(kPart) / sizeof(kPart[0]));
    return kPart;
}

}  // namespace gl
