/*
 * pt_tracer_qt.h -- the reference's Qt / glm signatures of class PathTracer
 * (PathTracerLib/src/PathTracer.h:30-152) over the plain-type facade of pt_tracer.h.
 *
 * Header-only and optional: include it only where Qt (QColor, QRect, QString) is available; the
 * glm overloads are compiled when <glm/vec3.hpp> is found. The Application can then keep its call
 * sites (OpenGLViewportWidget.cpp) unchanged:
 *   - configure(w, h, pbo0, pbo1, proceduralInstances, meshParams): the two GL pixel-buffer names
 *     are recorded (glPbo(slot)); frames are published into library-owned device buffers, which the
 *     caller copies into its PBOs (readDisplay, or HIP-GL interop on the GL thread);
 *   - ProceduralInstance and MeshSceneBuildParams keep the reference's field names and defaults
 *     (ProceduralTypes.h:80-86, MeshSceneContent.h:9-13);
 *   - setClearColor(QColor) ignores an invalid colour, as the reference does;
 *   - setEnvironmentHdrPath(QString) loads a Radiance .hdr file (ptEnvironmentLoadHdr) and an
 *     empty path returns to the solid background.
 */
#ifndef PT_TRACER_QT_H
#define PT_TRACER_QT_H

#include "pt_tracer.h"

#include <QColor>
#include <QRect>
#include <QString>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#if defined(__has_include)
#if __has_include(<glm/vec3.hpp>)
#include <glm/vec3.hpp>
#define PT_TRACER_QT_HAS_GLM 1
#endif
#endif

namespace ptqt {

struct ProceduralInstance {
    std::string commandString;
    std::size_t iterations = 0;
    Vec3 translation{0.0f, 0.0f, 0.0f};
    Vec3 rotationDeg{0.0f, 0.0f, 0.0f};
};

struct MeshSceneBuildParams {
    int circularSegments = 64;
    float creaseAngleDeg = 50.0f;
};

class PathTracer : public ::PathTracer {
public:
    bool configure(int width, int height, uint32_t pbo0, uint32_t pbo1,
                   const std::vector<ProceduralInstance>& proceduralInstances = {},
                   const MeshSceneBuildParams& meshParams = {})
    {
        m_pbo[0] = pbo0;
        m_pbo[1] = pbo1;
        return ::PathTracer::configure(width, height, plain(proceduralInstances), plain(meshParams));
    }

    bool rebuildMeshScene(const std::vector<ProceduralInstance>& proceduralInstances = {},
                          const MeshSceneBuildParams& meshParams = {})
    {
        return ::PathTracer::rebuildMeshScene(plain(proceduralInstances), plain(meshParams));
    }

    uint32_t glPbo(int slot) const { return slot == 0 || slot == 1 ? m_pbo[slot] : 0u; }

    void setClearColor(const QColor& color)
    {
        if (color.isValid()) {
            ::PathTracer::setClearColor(static_cast<float>(color.redF()), static_cast<float>(color.greenF()),
                                        static_cast<float>(color.blueF()));
        }
    }

    QRect renderRegion() const
    {
        const PtRegion r = ::PathTracer::renderRegion();
        return QRect(QPoint(r.minX, r.minY), QPoint(r.maxX, r.maxY));
    }

    void setEnvironmentHdrPath(const QString& path)
    {
        const QString p = path.trimmed();
        if (p == m_hdrPath) {
            return;
        }
        m_hdrPath = p;
        if (p.isEmpty()) {
            (void)setEnvironmentImage(nullptr, 0, 0);
            return;
        }
        const QByteArray utf8 = p.toUtf8();
        int w = 0, h = 0;
        float* rgb = ptEnvironmentLoadHdr(utf8.constData(), &w, &h);
        if (rgb != nullptr) {
            (void)setEnvironmentImage(rgb, w, h);
            ptEnvironmentFreeImage(rgb);
        }
    }
    QString environmentHdrPath() const { return m_hdrPath; }

#ifdef PT_TRACER_QT_HAS_GLM
    void setFocusPoint(const glm::vec3& p) { ::PathTracer::setFocusPoint(Vec3{p.x, p.y, p.z}); }
    bool pickSurface(const glm::vec3& ro, const glm::vec3& rd, glm::vec3* hitPoint) const
    {
        Vec3 hit{};
        if (!::PathTracer::pickSurface(Vec3{ro.x, ro.y, ro.z}, Vec3{rd.x, rd.y, rd.z}, &hit)) {
            return false;
        }
        if (hitPoint != nullptr) {
            *hitPoint = glm::vec3(hit.x, hit.y, hit.z);
        }
        return true;
    }
#endif
    using ::PathTracer::pickSurface;
    using ::PathTracer::setFocusPoint;

private:
    static std::vector<PtProceduralInstance> plain(const std::vector<ProceduralInstance>& in)
    {
        std::vector<PtProceduralInstance> out;
        out.reserve(in.size());
        for (const ProceduralInstance& i : in) {
            out.push_back(PtProceduralInstance{i.commandString.c_str(), static_cast<uint64_t>(i.iterations),
                                               i.translation, i.rotationDeg});
        }
        return out;
    }
    static PtMeshSceneBuildParams plain(const MeshSceneBuildParams& p)
    {
        return PtMeshSceneBuildParams{p.circularSegments, p.creaseAngleDeg};
    }

    uint32_t m_pbo[2] = {0u, 0u};
    QString m_hdrPath;
};

} // namespace ptqt

#endif /* PT_TRACER_QT_H */
