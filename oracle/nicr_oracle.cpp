// nicr_oracle.cpp -- CPU restatement of the reference Ni-Cr phase-field time step.
//
// TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP library is
// compared against. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it; the product library never links or calls it.
//
// It restates, operation for operation in IEEE binary64 (no FMA contraction,
// see oracle/Makefile), the members of `class Model` that form the per-step path
// of the two reference programs:
//   dim 2 = mpi2d/NiCr2D_grains.cpp          (cited as 2D:line)
//   dim 3 = mpi3d-slab/NiCr3D_grains.cpp     (cited as SL:line)
// Budgets and sums are long double, as in the reference (2D:193, SL:217).
//
// Storage keeps the reference's one-cell halo so that ghosts() is a direct
// restatement: 2D index i*(ny+2)+j (2D:149), 3D ((i*(ny+2)+j)*(nz+2)+k (SL:165-167).
// Initial fields (Voronoi mask, eta, center seed) are supplied by the caller
// through ora_set_field, exactly as a test would poke Model's public fields.
//
// Pinning: every function here is checked bit-for-bit against the reference
// itself compiled from /root/reference (oracle/_ref, see oracle/Makefile) by
// tests/test_oracle_vs_reference.py, and against the reference's closed-form
// self-test values (tests/test_oracle_closed_forms.py).
#include "../include/nicr.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <memory>
#include <sstream>
#include <iomanip>
#include <string>
#include <vector>

namespace {

// Physics constants (2D:65-71, SL:73-79).
struct Phys {
    static constexpr double Vm=1., W=1e4, ko=1e-9, kc=7.5e-9, kp=1.7e-10, L=0.2;
    static constexpr double bmo=2.21e-16, bmc=3.5436e-20, boo=2.21e-18, boc=3.5436e-21;
    static constexpr double ckO=2e-10, ckCr=2e-10, ksp=1.28e-7;
    static constexpr double bulkO=0., bulkCr=0.1986, layerO=0.0023, layerCr=0.182;
    static constexpr double oxideO=0.56300203, oxideCr=0.35677915;
};

// 2D:152, SL:171
double h(double p){return p*p*p*(10.-15.*p+6.*p*p);}
// 2D:79, SL:87
double hp(double p){return 30.*p*p*(p-1.)*(p-1.);}
// 2D:153, SL:172
double minimum_image(double d,double period){return d-period*std::round(d/period);}
// 2D:174-177, SL:198-201
double fmetal(double o,double cr){
    return (6000000.*o*o+583264.303945031*cr*cr+3200000.*o*cr
            -298432.*o-97045.64*cr-61732.2974223055)/Phys::Vm;
}
// 2D:178-181, SL:202-205
double foxide(double o,double cr){
    return (12000000.*o*o+7000000.*cr*cr-13800000.*o*cr
            -8880000.*o+2680000.*cr+1860990.)/Phys::Vm;
}
// 2D:182, SL:206
double mu_metal_o(double o,double cr){return (12000000.*o+3200000.*cr-298432.)/Phys::Vm;}

struct LBudget {long double nucO=0,nucCr=0,ckO=0,ckCr=0,clipO=0,clipCr=0,bcO=0,bcCr=0;long long nuclei=0,clipped=0;};

struct StepFailure { int status; int i,j,k,g; double o,cr,p,v; std::string msg; };

class Oracle {
public:
    nicr_config c;
    int dim,nx,ny,nz,nzs;   // nzs: halo-included Z extent (1 in 2D)
    std::size_t plane,total;
    std::vector<double> o,cr,p,on,crn,pn,muO,muCr,mo,mc,gb,s2,pairs;
    std::vector<std::vector<double>> eta,eta_new;
    LBudget b;
    std::string err;

    explicit Oracle(const nicr_config& cfg):c(cfg){
        dim=c.dim;nx=c.nx;ny=c.ny;nz=dim==3?c.nz:1;nzs=dim==3?nz+2:1;
        plane=std::size_t(ny+2)*nzs;total=std::size_t(nx+2)*plane;
        for(auto* f:{&o,&cr,&p,&on,&crn,&pn,&muO,&muCr,&mo,&mc,&gb,&s2,&pairs})f->assign(total,0.);
        eta.assign(std::size_t(c.n_eta),std::vector<double>(total,0.));
        eta_new.assign(std::size_t(c.n_eta),std::vector<double>(total,0.));
    }
    bool has_eta()const{return c.grain_evolution && c.grains && c.n_eta>0;}
    int klo()const{return dim==3?1:0;}
    int khi()const{return dim==3?nz:0;}
    std::size_t ix(int i,int j,int k)const{return (std::size_t(i)*(ny+2)+j)*nzs+k;}

    // 2D:154-169 / SL:173-191 (single rank: only the physical BC part runs).
    void ghosts(std::vector<double>& a)const{
        if(dim==2){
            for(int j=1;j<=ny;++j)a[ix(0,j,0)]=a[ix(1,j,0)];
            for(int j=1;j<=ny;++j)a[ix(nx+1,j,0)]=a[ix(nx,j,0)];
            for(int i=0;i<=nx+1;++i){a[ix(i,0,0)]=a[ix(i,ny,0)];a[ix(i,ny+1,0)]=a[ix(i,1,0)];}
            return;
        }
        for(int j=1;j<=ny;++j)for(int k=1;k<=nz;++k)a[ix(0,j,k)]=a[ix(1,j,k)];
        for(int j=1;j<=ny;++j)for(int k=1;k<=nz;++k)a[ix(nx+1,j,k)]=a[ix(nx,j,k)];
        for(int i=0;i<=nx+1;++i){
            for(int k=1;k<=nz;++k){a[ix(i,0,k)]=a[ix(i,ny,k)];a[ix(i,ny+1,k)]=a[ix(i,1,k)];}
            for(int j=0;j<=ny+1;++j){a[ix(i,j,0)]=a[ix(i,j,nz)];a[ix(i,j,nz+1)]=a[ix(i,j,1)];}
        }
    }
    // 2D:170-173, SL:192-196
    double lap(const std::vector<double>& a,int i,int j,int k)const{
        const double xy=(a[ix(i+1,j,k)]-2.*a[ix(i,j,k)]+a[ix(i-1,j,k)])/(c.dx*c.dx)
                       +(a[ix(i,j+1,k)]-2.*a[ix(i,j,k)]+a[ix(i,j-1,k)])/(c.dy*c.dy);
        if(dim==2)return xy;
        return xy+(a[ix(i,j,k+1)]-2.*a[ix(i,j,k)]+a[ix(i,j,k-1)])/(c.dz*c.dz);
    }
    // 2D:183-191, SL:207-215
    double mobility(double conc,double pp,double g,bool oxygen)const{
        if(oxygen && !c.oxygen_diffusion)return 0.;
        double hh=h(pp), bm=oxygen?Phys::bmo:Phys::bmc, bo=oxygen?Phys::boo:Phys::boc;
        const double gb_factor=oxygen?c.gb_factor_o:c.gb_factor_cr;
        const double active_gb=c.grains?g:0.;
        return conc*((1.-hh)*bm*(1.+(gb_factor-1.)*active_gb)+hh*bo);
    }
    // grain_local_deta / grain_local_dphi (2D:100-105, SL:108-113)
    double deta(double e,double S,double phi)const{
        return e*(c.eta_W*(S-(1.-h(phi))+2.*(S-e*e))+c.eta_A*h(phi));
    }
    double dphi(double S,double phi)const{
        return 0.5*hp(phi)*(c.eta_W*(S-(1.-h(phi)))+c.eta_A*S);
    }
    // 2D:328-340, SL:495-507
    void rebuild_eta_mask(){
        const double scale=c.eta_scale;
        for(int i=1;i<=nx;++i)for(int j=1;j<=ny;++j)for(int k=klo();k<=khi();++k){
            double S=0.,pr=0.;
            for(const auto& e:eta){const double q=e[ix(i,j,k)]*e[ix(i,j,k)];pr+=S*q;S+=q;}
            s2[ix(i,j,k)]=S;pairs[ix(i,j,k)]=pr;
            gb[ix(i,j,k)]=scale*std::min(1.,16.*pr);
        }
        ghosts(gb);
    }
    double eta_force(std::size_t g,int i,int j,int k)const{   // 2D:341-343, SL:508-510
        return deta(eta[g][ix(i,j,k)],s2[ix(i,j,k)],p[ix(i,j,k)])-c.eta_K*lap(eta[g],i,j,k);
    }
    double eta_phi_force(int i,int j,int k)const{              // 2D:344-346, SL:511-513
        return has_eta()?dphi(s2[ix(i,j,k)],p[ix(i,j,k)]):0.;
    }
    // 2D:408-411, SL:593-596
    void enforce_legacy_layer(){
        if(c.bc!=NICR_BC_CONCENTRATION || !c.oxygen_diffusion)return;
        for(int j=1;j<=ny;++j)for(int k=klo();k<=khi();++k){
            b.bcO+=c.layer_o-o[ix(1,j,k)];o[ix(1,j,k)]=c.layer_o;   // layer_o: Physics::layerO by default
        }
    }
    // 2D:412-434, SL:597-620 (the SL frozen-eta cache is bitwise equal to a rebuild,
    // SL:346-349 and regression_driver.cpp:60-64, so the mask is always rebuilt here)
    void refresh(){
        ghosts(o);ghosts(cr);ghosts(p);
        if(has_eta()){
            for(auto& e:eta)ghosts(e);
            rebuild_eta_mask();
        }
        for(int i=1;i<=nx;++i)for(int j=1;j<=ny;++j)for(int k=klo();k<=khi();++k){
            const std::size_t q=ix(i,j,k);
            double hh=h(p[q]);
            muO[q]=(1.-hh)*mu_metal_o(o[q],cr[q])
                +hh*(24000000.*o[q]-13800000.*cr[q]-8880000.)/Phys::Vm
                -2.*Phys::ko*lap(o,i,j,k);
            muCr[q]=(1.-hh)*(1166528.60789006*cr[q]+3200000.*o[q]-97045.64)/Phys::Vm
                +hh*(14000000.*cr[q]-13800000.*o[q]+2680000.)/Phys::Vm
                -2.*Phys::kc*lap(cr,i,j,k);
        }
        ghosts(muO);ghosts(muCr);
        if(c.bc==NICR_BC_MU && c.oxygen_diffusion)
            for(int j=1;j<=ny;++j)for(int k=klo();k<=khi();++k)muO[ix(0,j,k)]=2.*c.mu_res-muO[ix(1,j,k)];
        const int k0=dim==3?0:0,k1=dim==3?nz+1:0;
        for(int i=0;i<=nx+1;++i)for(int j=0;j<=ny+1;++j)for(int k=k0;k<=k1;++k){
            const std::size_t q=ix(i,j,k);
            mo[q]=mobility(o[q],p[q],gb[q],true);
            mc[q]=mobility(cr[q],p[q],gb[q],false);
        }
    }
    // 2D:435-442, SL:621-628
    double left_flux(int j,int k)const{
        if(c.bc!=NICR_BC_MU || !c.oxygen_diffusion)return 0.;
        const double mf=mobility(c.surface_c_ref,p[ix(1,j,k)],gb[ix(1,j,k)],true);
        return 2.*mf*(c.mu_res-muO[ix(1,j,k)])/c.dx;
    }
    // 2D:443-450, SL:629-638
    double div_flux(const std::vector<double>& m,const std::vector<double>& mu,int i,int j,int k,bool oxygen)const{
        const double right=-.5*(m[ix(i,j,k)]+m[ix(i+1,j,k)])*(mu[ix(i+1,j,k)]-mu[ix(i,j,k)])/c.dx;
        double left=-.5*(m[ix(i-1,j,k)]+m[ix(i,j,k)])*(mu[ix(i,j,k)]-mu[ix(i-1,j,k)])/c.dx;
        if(oxygen && i==1 && c.bc==NICR_BC_MU)left=left_flux(j,k);
        const double top=-.5*(m[ix(i,j,k)]+m[ix(i,j+1,k)])*(mu[ix(i,j+1,k)]-mu[ix(i,j,k)])/c.dy;
        const double bottom=-.5*(m[ix(i,j-1,k)]+m[ix(i,j,k)])*(mu[ix(i,j,k)]-mu[ix(i,j-1,k)])/c.dy;
        if(dim==2)return (left-right)/c.dx+(bottom-top)/c.dy;
        const double front=-.5*(m[ix(i,j,k)]+m[ix(i,j,k+1)])*(mu[ix(i,j,k+1)]-mu[ix(i,j,k)])/c.dz;
        const double back=-.5*(m[ix(i,j,k-1)]+m[ix(i,j,k)])*(mu[ix(i,j,k)]-mu[ix(i,j,k-1)])/c.dz;
        return (left-right)/c.dx+(bottom-top)/c.dy+(back-front)/c.dz;
    }
    // 2D:347-359, SL:514-526
    bool eta_trial(StepFailure& f){
        for(std::size_t g=0;g<eta.size();++g)
            for(int i=1;i<=nx;++i)for(int j=1;j<=ny;++j)for(int k=klo();k<=khi();++k){
                const double v=eta[g][ix(i,j,k)]-c.dt*c.eta_L*eta_force(g,i,j,k);
                if(!std::isfinite(v)||v < -1.e-12||v > 1.+1.e-12){
                    f.status=NICR_STEP_ETA_RANGE;f.i=i;f.j=j;f.k=dim==3?k:0;f.g=int(g)+1;f.v=v;
                    return false;
                }
                eta_new[g][ix(i,j,k)]=v;
            }
        return true;
    }
    // 2D:551-597, SL:761-826. Returns NICR_STEP_*; on failure the fields keep
    // their pre-step values and the budget is restored (the reference's caller
    // restores it in 2D, 2D:1224-1225; SL restores it and the inlet, SL:812-820).
    // The 2D program leaves the inlet layer reset in place on failure (2D:552).
    int step(StepFailure& f){
        const LBudget before=b;
        std::vector<double> inlet_before;
        const bool reset_inlet=c.bc==NICR_BC_CONCENTRATION && c.oxygen_diffusion;
        if(dim==3 && reset_inlet)
            for(int j=1;j<=ny;++j)for(int k=klo();k<=khi();++k)inlet_before.push_back(o[ix(1,j,k)]);
        enforce_legacy_layer();refresh();
        for(int j=1;j<=ny;++j)for(int k=klo();k<=khi();++k)b.bcO+=c.dt*left_flux(j,k)/c.dx;
        f=StepFailure{};f.status=NICR_STEP_OK;
        bool ok=true;
        for(int i=1;i<=nx && ok;++i)for(int j=1;j<=ny && ok;++j)for(int k=klo();k<=khi() && ok;++k){
            const std::size_t q=ix(i,j,k);
            double oo=o[q],cc=cr[q],pp=p[q];
            on[q]=c.oxygen_diffusion?oo+c.dt*div_flux(mo,muO,i,j,k,true):oo;
            crn[q]=cc+c.dt*div_flux(mc,muCr,i,j,k,false);
            const double hpp=30.*pp*pp*(pp-1.)*(pp-1.);
            double fp=Phys::W*(4.*pp*pp*pp-6.*pp*pp+2.*pp)
                      +hpp*(foxide(oo,cc)-fmetal(oo,cc))-2.*Phys::kp*lap(p,i,j,k);
            if(has_eta())fp+=eta_phi_force(i,j,k);
            pn[q]=pp-c.dt*Phys::L*fp;
            if(c.ck==NICR_CK_LEGACY){
                double dO=-Phys::ckO*((Phys::oxideO-on[q])*h(pn[q])-(Phys::oxideO-oo)*h(pp))/c.dt;
                double dCr=-Phys::ckCr*((Phys::oxideCr-crn[q])*h(pn[q])-(Phys::oxideCr-cc)*h(pp))/c.dt;
                on[q]+=dO;crn[q]+=dCr;b.ckO+=dO;b.ckCr+=dCr;
            }
            if(!std::isfinite(on[q]) || !std::isfinite(crn[q]) || !std::isfinite(pn[q])){
                f.status=NICR_STEP_NONFINITE;f.i=i;f.j=j;f.k=dim==3?k:0;ok=false;break;
            }
            if(c.bounds==NICR_BOUNDS_LEGACY){
                double oc=std::clamp(on[q],0.,1.),ccr=std::clamp(crn[q],0.,1.);
                b.clipO+=oc-on[q];b.clipCr+=ccr-crn[q];
                if(oc!=on[q]||ccr!=crn[q])++b.clipped;
                on[q]=oc;crn[q]=ccr;
                if(pn[q]>.99999)pn[q]=1.;
                if(pn[q]<1.e-5)pn[q]=0.;
            }else{
                if(on[q]<-1.e-12 || crn[q]<-1.e-12 || on[q]+crn[q]>1.+1.e-12 ||
                   pn[q]<-1.e-12 || pn[q]>1.+1.e-12){
                    f.status=NICR_STEP_OUT_OF_BOUNDS;f.i=i;f.j=j;f.k=dim==3?k:0;
                    f.o=on[q];f.cr=crn[q];f.p=pn[q];ok=false;break;
                }
            }
        }
        const bool run_eta=has_eta() && (dim==2 || c.eta_L>0.);
        if(ok && run_eta)ok=eta_trial(f);
        if(!ok){
            b=before;
            if(dim==3 && reset_inlet){
                std::size_t q=0;
                for(int j=1;j<=ny;++j)for(int k=klo();k<=khi();++k)o[ix(1,j,k)]=inlet_before[q++];
            }
            return f.status;
        }
        o.swap(on);cr.swap(crn);p.swap(pn);
        if(run_eta)for(std::size_t g=0;g<eta.size();++g)eta[g].swap(eta_new[g]);
        return NICR_STEP_OK;
    }
    // ---- Ksp insertion ----
    // 3D bounded periodic neighbourhood (SL:220-231).
    struct Axis { int count=0; int lo[2]={0,0},hi[2]={0,0}; };
    static int bounded_reach(double radius,int n){return radius>=n?n:int(std::ceil(radius));}
    static Axis periodic_neighborhood(int center,double radius,int n){
        const int r=bounded_reach(radius,n);Axis a;
        if(2LL*r+1>=n){a.count=1;a.lo[0]=1;a.hi[0]=n;return a;}
        const long long lo=static_cast<long long>(center)-r,hi=static_cast<long long>(center)+r;
        if(lo<1){a.count=2;a.lo[0]=1;a.lo[1]=int(lo+n);a.hi[0]=int(hi);a.hi[1]=n;}
        else if(hi>n){a.count=2;a.lo[0]=1;a.lo[1]=int(lo);a.hi[0]=int(hi-n);a.hi[1]=n;}
        else{a.count=1;a.lo[0]=int(lo);a.hi[0]=int(hi);}
        return a;
    }
    // 2D:493-502 (full plane) / SL:687-703 (bounded neighbourhood, same predicate)
    bool valid_nucleus(int x,int y,int z)const{
        const double r=c.radius_grid;
        if(x-r<5. || x+r>double(nx-3))return false;
        if(dim==2){
            int sr=int(std::ceil(r+c.exclusion_gap_grid));
            for(int i=std::max(1,x-sr);i<=std::min(nx,x+sr);++i)for(int j=1;j<=ny;++j){
                if(p[ix(i,j,0)]>.01 && std::hypot(double(i-x),minimum_image(j-y,ny))<r+c.exclusion_gap_grid)return false;
            }
            return true;
        }
        const double limit=r+c.exclusion_gap_grid;
        const auto ys=periodic_neighborhood(y,limit,ny),zs=periodic_neighborhood(z,limit,nz);
        const int sr=bounded_reach(limit,nx);
        for(int i=int(std::max(1LL,static_cast<long long>(x)-sr));
            i<=int(std::min(static_cast<long long>(nx),static_cast<long long>(x)+sr));++i)
        for(int jy=0;jy<ys.count;++jy)for(int j=ys.lo[jy];j<=ys.hi[jy];++j)
        for(int kz=0;kz<zs.count;++kz)for(int k=zs.lo[kz];k<=zs.hi[kz];++k){
            if(p[ix(i,j,k)]>.01 && std::sqrt(std::pow(double(i-x),2)+std::pow(minimum_image(j-y,ny),2)+std::pow(minimum_image(k-z,nz),2))<r+c.exclusion_gap_grid)return false;
        }
        return true;
    }
    void stamp_cell(std::size_t q,int& changed){
        b.nucO+=Phys::oxideO-o[q];b.nucCr+=Phys::oxideCr-cr[q];
        o[q]=Phys::oxideO;cr[q]=Phys::oxideCr;p[q]=1.;++changed;
        for(auto& e:eta)e[q]=0.;
    }
    // initialize_center_oxide, 2D:517-542 / SL:725-751: returns the cell count,
    // -1 for a radius at or above (min side - 1)/2, -2 with a nucleus already
    // recorded (the reference throws); 0 cells leave nuclei unchanged.
    int center_seed(){
        const double x=0.5*(double(nx)+1.),y=0.5*(double(ny)+1.),z=0.5*(double(nz)+1.);
        const double r=c.radius_grid;
        const int nmin=dim==2?std::min(nx,ny):std::min({nx,ny,nz});
        if(r>=0.5*(double(nmin)-1.))return -1;
        if(b.nuclei!=0)return -2;
        int cells=0;
        for(int i=1;i<=nx;++i)for(int j=1;j<=ny;++j)for(int k=klo();k<=khi();++k){
            const double dx=i-x,dy=j-y;
            double d2=dx*dx+dy*dy;
            if(dim==3){const double dz=k-z;d2=dx*dx+dy*dy+dz*dz;}
            if(d2<=r*r)stamp_cell(ix(i,j,k),cells);
        }
        if(cells)++b.nuclei;
        return cells;
    }
    // 2D:503-516 / SL:704-724
    int stamp(int x,int y,int z){
        int changed=0;
        if(dim==2){
            int rr=int(std::ceil(c.radius_grid));
            for(int i=std::max(1,x-rr);i<=std::min(nx,x+rr);++i)for(int j=1;j<=ny;++j){
                double dy=minimum_image(j-y,ny);
                if((i-x)*(i-x)+dy*dy<=c.radius_grid*c.radius_grid && p[ix(i,j,0)]<.01)stamp_cell(ix(i,j,0),changed);
            }
        }else{
            const int rr=bounded_reach(c.radius_grid,nx);
            const auto ys=periodic_neighborhood(y,c.radius_grid,ny),zs=periodic_neighborhood(z,c.radius_grid,nz);
            for(int i=int(std::max(1LL,static_cast<long long>(x)-rr));
                i<=int(std::min(static_cast<long long>(nx),static_cast<long long>(x)+rr));++i)
            for(int jy=0;jy<ys.count;++jy)for(int j=ys.lo[jy];j<=ys.hi[jy];++j)
            for(int kz=0;kz<zs.count;++kz)for(int k=zs.lo[kz];k<=zs.hi[kz];++k){
                double dy=minimum_image(j-y,ny),dz=minimum_image(k-z,nz);
                if((i-x)*(i-x)+dy*dy+dz*dz<=c.radius_grid*c.radius_grid && p[ix(i,j,k)]<.01)stamp_cell(ix(i,j,k),changed);
            }
        }
        if(changed)++b.nuclei;
        return changed;
    }
    // 2D:543-550, SL:752-759
    int nucleate(){
        if(c.nucleation!=NICR_NUC_KSP)return 0;
        int cells=0;
        for(int i=1;i<=nx;++i)for(int j=1;j<=ny;++j)for(int k=klo();k<=khi();++k){
            const std::size_t q=ix(i,j,k);
            if(p[q]<=.01 && o[q]*cr[q]>Phys::ksp && valid_nucleus(i,j,dim==3?k:0))cells+=stamp(i,j,dim==3?k:0);
        }
        return cells;
    }
    // 2D:229-232, SL:389-392
    long double sum(const std::vector<double>& a)const{
        long double v=0;for(int i=1;i<=nx;++i)for(int j=1;j<=ny;++j)for(int k=klo();k<=khi();++k)v+=a[ix(i,j,k)];
        return v;
    }
    std::vector<double>* field(int id){
        switch(id){
        case NICR_FIELD_O:return &o;case NICR_FIELD_CR:return &cr;case NICR_FIELD_PHI:return &p;
        case NICR_FIELD_GB:return &gb;case NICR_FIELD_MU_O:return &muO;case NICR_FIELD_MU_CR:return &muCr;
        case NICR_FIELD_M_O:return &mo;case NICR_FIELD_M_CR:return &mc;
        case NICR_FIELD_ETA_S2:return &s2;case NICR_FIELD_ETA_PAIRS:return &pairs;
        default:
            if(id>=NICR_FIELD_ETA0 && id<NICR_FIELD_ETA0+int(eta.size()))return &eta[std::size_t(id-NICR_FIELD_ETA0)];
            return nullptr;
        }
    }
};

thread_local std::string g_error;
void split(long double v,double out[2]){out[0]=double(v);out[1]=double(v-(long double)out[0]);}
long double join(const double in[2]){return (long double)in[0]+(long double)in[1];}

}  // namespace

extern "C" {

const char* ora_last_error(void){return g_error.c_str();}

int ora_create(const nicr_config* cfg,void** out){
    if(!cfg||!out||(cfg->dim!=2&&cfg->dim!=3)||cfg->nx<1||cfg->ny<1||(cfg->dim==3&&cfg->nz<1)){g_error="invalid config";return NICR_E_INVALID;}
    try{*out=new Oracle(*cfg);}catch(const std::exception& e){g_error=e.what();return NICR_E_INVALID;}
    return NICR_OK;
}
int ora_destroy(void* h){delete static_cast<Oracle*>(h);return NICR_OK;}

// Interior copy in/out (host layout documented in include/nicr.h), then ghosts.
int ora_upload_field(void* h,int id,const double* src){
    auto* m=static_cast<Oracle*>(h);auto* f=m->field(id);
    if(!f){g_error="unknown field";return NICR_E_INVALID;}
    std::size_t q=0;
    for(int i=1;i<=m->nx;++i)for(int j=1;j<=m->ny;++j)for(int k=m->klo();k<=m->khi();++k)(*f)[m->ix(i,j,k)]=src[q++];
    m->ghosts(*f);
    return NICR_OK;
}
int ora_download_field(void* h,int id,double* dst){
    auto* m=static_cast<Oracle*>(h);auto* f=m->field(id);
    if(!f){g_error="unknown field";return NICR_E_INVALID;}
    std::size_t q=0;
    for(int i=1;i<=m->nx;++i)for(int j=1;j<=m->ny;++j)for(int k=m->klo();k<=m->khi();++k)dst[q++]=(*f)[m->ix(i,j,k)];
    return NICR_OK;
}
int ora_set_budget(void* h,const nicr_budget* b){
    auto* m=static_cast<Oracle*>(h);
    m->b.nucO=join(b->nucO);m->b.nucCr=join(b->nucCr);m->b.ckO=join(b->ckO);m->b.ckCr=join(b->ckCr);
    m->b.clipO=join(b->clipO);m->b.clipCr=join(b->clipCr);m->b.bcO=join(b->bcO);m->b.bcCr=join(b->bcCr);
    m->b.nuclei=b->nuclei;m->b.clipped=b->clipped;return NICR_OK;
}
int ora_get_budget(void* h,nicr_budget* b){
    auto* m=static_cast<Oracle*>(h);
    split(m->b.nucO,b->nucO);split(m->b.nucCr,b->nucCr);split(m->b.ckO,b->ckO);split(m->b.ckCr,b->ckCr);
    split(m->b.clipO,b->clipO);split(m->b.clipCr,b->clipCr);split(m->b.bcO,b->bcO);split(m->b.bcCr,b->bcCr);
    b->nuclei=m->b.nuclei;b->clipped=m->b.clipped;return NICR_OK;
}
int ora_refresh(void* h){static_cast<Oracle*>(h)->refresh();return NICR_OK;}
int ora_step(void* h,int nsteps,nicr_step_report* rep){
    auto* m=static_cast<Oracle*>(h);
    nicr_step_report r{};r.status=NICR_STEP_OK;
    for(int n=0;n<nsteps;++n){
        StepFailure f{};
        const int st=m->step(f);
        if(st!=NICR_STEP_OK){
            r.status=st;r.i=f.i;r.j=f.j;r.k=f.k;r.grain=f.g;r.o=f.o;r.cr=f.cr;r.phi=f.p;r.eta_value=f.v;
            break;
        }
        ++r.steps_done;
    }
    if(rep)*rep=r;
    return NICR_OK;
}
int ora_nucleate(void* h,int* cells){int n=static_cast<Oracle*>(h)->nucleate();if(cells)*cells=n;return NICR_OK;}
int ora_center_seed(void* h,int* cells){
    Oracle* m=static_cast<Oracle*>(h);
    const int n=m->center_seed();
    if(cells)*cells=n>0?n:0;
    if(n==-1){g_error=m->dim==3?"Center seed is too large: --radius-grid must be less than (min(nx,ny,nz)-1)/2."
                                :"Center seed is too large: --radius-grid must be less than (min(nx,ny)-1)/2.";return NICR_E_INVALID;}
    if(n==-2){g_error="The center oxide has already been initialized.";return NICR_E_STATE;}
    if(n==0){g_error="Center seed covers no cell centres: increase --radius-grid.";return NICR_E_INVALID;}
    return NICR_OK;
}
int ora_stamp(void* h,int x,int y,int z,int* changed){int n=static_cast<Oracle*>(h)->stamp(x,y,z);if(changed)*changed=n;return NICR_OK;}
// Model::sum in long double, returned split as hi/lo binary64.
int ora_sum(void* h,int id,double out[2]){
    auto* m=static_cast<Oracle*>(h);auto* f=m->field(id);
    if(!f){g_error="unknown field";return NICR_E_INVALID;}
    split(m->sum(*f),out);return NICR_OK;
}
// valid_nucleus exposed for the insertion tests (2D:493-502, SL:687-703).
int ora_valid_nucleus(void* h,int x,int y,int z){return static_cast<Oracle*>(h)->valid_nucleus(x,y,z)?1:0;}
// div_flux / lap at one cell (after ora_refresh), for the closed-form tests.
double ora_div_flux(void* h,int species,int i,int j,int k){
    auto* m=static_cast<Oracle*>(h);
    return species==0?m->div_flux(m->mo,m->muO,i,j,k,true):m->div_flux(m->mc,m->muCr,i,j,k,false);
}
double ora_lap(void* h,int id,int i,int j,int k){
    auto* m=static_cast<Oracle*>(h);auto* f=m->field(id);return f?m->lap(*f,i,j,k):NAN;
}
double ora_mobility(void* h,double conc,double p,double gb,int oxygen){
    return static_cast<Oracle*>(h)->mobility(conc,p,gb,oxygen!=0);
}
double ora_mu_metal_o(double o,double cr){return mu_metal_o(o,cr);}
// One-thread wall-clock benchmark helper: nsteps without report bookkeeping.
int ora_run(void* h,int nsteps){
    auto* m=static_cast<Oracle*>(h);
    for(int n=0;n<nsteps;++n){StepFailure f{};if(m->step(f)!=NICR_STEP_OK)return n;}
    return nsteps;
}

}  // extern "C"
